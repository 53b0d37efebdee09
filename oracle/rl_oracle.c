/*
 * CPU FP64 restatement of RADLite's circular-image line raytrace
 * (calc_line_spectrum and everything below it).  TEST INFRASTRUCTURE
 * ONLY: the checker for the HIP engine and the timed CPU baseline.
 *
 * Every routine follows a reference routine, cited as file:line under
 * RADLITE/.  Arithmetic is written in the reference's evaluation order
 * (Fortran evaluates a*b*c as (a*b)*c, a+b+c as (a+b)+c; ``x**2`` is
 * x*x) and uses the same libm, so on x86-64 without FMA the results are
 * meant to be bit-identical to the reference build in oracle/refbuild.
 * Build with -O2 -ffp-contract=off (oracle/Makefile).
 *
 * Parity status: pinned against the compiled reference's per-ray,
 * per-channel intensities and spectra (tests/golden) and against closed
 * forms (tests/test_oracle_*.py).  One reference behaviour is undefined
 * and is given a definition here: make_trajectory_c reads r_s(isrt-1)
 * with isrt=1 (telescope.F:3466-3468) for rays whose impact parameter
 * lies in the outermost radial shell, i.e. an uninitialised stack slot.
 * This restatement uses the value the reference build was observed to
 * produce for that slot (see RS0_FILL in orc_trajectory's callers).
 */
#include "rl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TELESC_EPS 1.0e-10              /* configure.h:46                 */
#define PICONST 3.1415926535897932385   /* main.h:452                     */
#define RAYEXPT 100                     /* common_telescope.h:4-6         */
#define RAYADPT 4
#define RAYRNPT 4
#define LEVTHRES 1.0e-3                 /* configure.h:52                 */

/* ------------------------------------------------------------------ */
/* Numerical Recipes hunt (nrecip.F:157-204), 1-based xx[1..n].        */
/* With jlo outside [1,n] this is a pure bisection; for x strictly     */
/* between table values the result does not depend on the guess.       */
/* ------------------------------------------------------------------ */
int orc_hunt(const double *xx, int n, double x, int jlo)
{
    int ascnd = xx[n] > xx[1];
    int jhi, inc, jm;
    if (jlo <= 0 || jlo > n) {
        jlo = 0;
        jhi = n + 1;
    } else {
        inc = 1;
        if ((x >= xx[jlo]) == ascnd) {
            for (;;) {
                jhi = jlo + inc;
                if (jhi > n) { jhi = n + 1; break; }
                if ((x >= xx[jhi]) == ascnd) { jlo = jhi; inc += inc; }
                else break;
            }
        } else {
            jhi = jlo;
            for (;;) {
                jlo = jhi - inc;
                if (jlo < 1) { jlo = 0; break; }
                if ((x < xx[jlo]) == ascnd) { jhi = jlo; inc += inc; }
                else break;
            }
        }
    }
    while (jhi - jlo != 1) {
        jm = (jhi + jlo) / 2;
        if ((x > xx[jm]) == ascnd) jlo = jm; else jhi = jm;
    }
    return jlo;
}

/* ------------------------------------------------------------------ */
/* Grid (grid.F:711-819 radius, grid.F:1098-1176 theta).               */
/* ------------------------------------------------------------------ */
#define RC(g, i) ((g)->rc[(i) + 1])
#define TC(g, i) ((g)->tc[(i) + 1])

int orc_grid_build(orc_grid *g, int nr, const double *r, int nth_half,
                   const double *theta_half)
{
    int i, ny = 2 * nth_half;
    if (nr < 3 || nth_half < 2) return 13;
    g->nr = nr;
    g->nth = ny;
    g->rc = (double *)calloc(nr + 4, sizeof(double));
    g->tc = (double *)calloc(ny + 4, sizeof(double));
    for (i = 1; i <= nr; i++) RC(g, i) = r[i - 1];
    /* grid.F:753-760 ghosts */
    RC(g, 0) = RC(g, 1) * RC(g, 1) / RC(g, 2);
    RC(g, -1) = RC(g, 0) * RC(g, 0) / RC(g, 1);
    RC(g, nr + 1) = RC(g, nr) * RC(g, nr) / RC(g, nr - 1);
    RC(g, nr + 2) = RC(g, nr + 1) * RC(g, nr + 1) / RC(g, nr);
    for (i = 1; i <= nth_half; i++) TC(g, i) = theta_half[i - 1];
    /* grid.F:1158-1163 mirror with pi = 3.14159265359d0 */
    for (i = 1; i <= nth_half; i++)
        TC(g, 2 * nth_half + 1 - i) = 3.14159265359 - TC(g, i);
    /* grid.F:1170-1176: outer ghosts use a single-precision 2*pi */
    {
        const float twopi_sp = 2 * 3.1415926f;
        TC(g, 0) = -TC(g, 1);
        TC(g, -1) = -TC(g, 2);
        TC(g, ny + 1) = (double)twopi_sp - TC(g, ny);
        TC(g, ny + 2) = (double)twopi_sp - TC(g, ny - 1);
    }
    return 0;
}

void orc_grid_free(orc_grid *g)
{
    free(g->rc);
    free(g->tc);
    g->rc = g->tc = NULL;
}

/* make_index theta re-index (interpol.F:87-100) */
static int ridx_it(int it, int ny)
{
    int v = it;
    if (v < 1) v = 1 - v;
    if (v > ny) v = 2 * ny + 1 - v;
    if (v > ny / 2) v = ny + 1 - v;
    return v;
}

/* ------------------------------------------------------------------ */
/* Rays (setup_rays_circular, telescope.F:715-1191, imethod=1,         */
/* nrref=10 as hard-wired at main.F:37-38; ring edges from             */
/* make_image_circular, telescope.F:434-488).                          */
/* ------------------------------------------------------------------ */
int orc_setup_rays(orc_rays *rays, const orc_grid *g, double anginf,
                   int nphi, int nrext, int dbdr, double rstar)
{
    const int nrref = 10;
    const double epsxyz = 1.0e2 * TELESC_EPS;
    const double epsrrr = 1.0e3 * TELESC_EPS;
    int nr = g->nr, nrrextra = nrext < 0 ? -nrext : nrext;
    int ix, iys, iins, ir, iradius, nrings;
    double theta0, sinth0, dphi, phi, r_c;
    double refdum1, refdum2;
    double *xhat, *zhat;

    if (nrrextra == 0) return 13;               /* main.F:628-635     */
    if (dbdr < 1) return 13;
    if (fabs(anginf) < 1.0e-1)                   /* telescope.F:818-827 */
        anginf = (double)0.1f * fabs(anginf) / anginf;
    if (rstar > RC(g, 1)) return 91991;

    theta0 = anginf + 1.0e-4;                    /* :863               */
    sinth0 = sin(theta0);
    dphi = 6.28318530718 / (1.0 * nphi);
    phi = 0.5 * dphi;
    xhat = (double *)malloc(sizeof(double) * nphi);
    zhat = (double *)malloc(sizeof(double) * nphi);
    for (iys = 0; iys < nphi; iys++) {           /* :887-903           */
        double zh02, dum;
        zhat[iys] = -sin(phi) / sinth0;
        zh02 = zhat[iys] * zhat[iys];
        dum = 1.0 - zh02 * sinth0 * sinth0;
        dum = dum + epsxyz;
        if (dum < 0.0) { free(xhat); free(zhat); return 13; }
        xhat[iys] = cos(phi) > 0.0 ? sqrt(dum) : -sqrt(dum);
        phi = phi + dphi;
    }
    nrings = nrrextra + nrref + (nr - 1) * dbdr;
    rays->nrings = nrings;
    rays->nphi = nphi;
    rays->nrays = 1 + nrings * nphi;
    rays->theta0 = theta0;
    rays->x0 = (double *)calloc(rays->nrays, sizeof(double));
    rays->z0 = (double *)calloc(rays->nrays, sizeof(double));
    rays->ring_r = (double *)calloc(nrings + 1, sizeof(double));
    rays->ring_ri = (double *)calloc(nrings + 2, sizeof(double));

    ir = 1;                                       /* central ray: 0,0   */
    iradius = 1;
    refdum1 = 0.5;
    refdum2 = 0.0;
    for (ix = 1; ix <= nrrextra + nrref; ix++) {  /* :1004-1056         */
        if (ix <= nrrextra)
            r_c = ((double)(ix - 1) * (RC(g, 1) - rstar) / nrrextra) + rstar;
        else {
            refdum2 = refdum2 + refdum1;
            refdum1 = refdum1 / 2;
            r_c = (((double)nrrextra + refdum2) * (RC(g, 1) - rstar) /
                   ((double)nrrextra + 1.0)) + rstar;
        }
        rays->ring_r[iradius++] = r_c;
        for (iys = 0; iys < nphi; iys++, ir++) {
            rays->x0[ir] = r_c * xhat[iys];
            rays->z0[ir] = r_c * zhat[iys];
        }
    }
    for (ix = 1; ix <= nr - 1; ix++) {            /* :1066-1138         */
        r_c = RC(g, ix) * (1.0 + epsrrr);
        rays->ring_r[iradius++] = r_c;
        for (iys = 0; iys < nphi; iys++, ir++) {
            rays->x0[ir] = r_c * xhat[iys];
            rays->z0[ir] = r_c * zhat[iys];
        }
        if (dbdr > 1) {
            double dr = (RC(g, ix + 1) - RC(g, ix)) / (1.0 * dbdr);
            for (iins = 1; iins <= dbdr - 1; iins++) {
                r_c = RC(g, ix) + iins * dr;
                rays->ring_r[iradius++] = r_c;
                for (iys = 0; iys < nphi; iys++, ir++) {
                    rays->x0[ir] = r_c * xhat[iys];
                    rays->z0[ir] = r_c * zhat[iys];
                }
            }
        }
    }
    free(xhat);
    free(zhat);

    /* ring edges, telescope.F:443-488 */
    {
        double *rr = rays->ring_r, *ri = rays->ring_ri;
        int k;
        ri[0] = 0.0;
        for (k = 1; k <= nrings; k++) ri[k] = 0.5 * (rr[k] + rr[k - 1]);
        ri[nrings + 1] = RC(g, nr);
        for (k = 1; k <= nrings - 1; k++)
            if (rr[k] - ri[k] > 2 * (ri[k + 1] - rr[k]))
                ri[k] = rr[k] - 2 * (ri[k + 1] - rr[k]);
        if (ri[1] < rstar) {
            if (rr[1] < rstar) return 1;
            ri[1] = rstar;
        }
    }
    return 0;
}

void orc_rays_free(orc_rays *r)
{
    free(r->x0); free(r->z0); free(r->ring_r); free(r->ring_ri);
    memset(r, 0, sizeof(*r));
}

/* ------------------------------------------------------------------ */
/* Trajectory (make_trajectory_c, telescope.F:2787-3720), send=1e30,   */
/* istar=0.  rs0_fill is the value used for the reference's            */
/* uninitialised r_s(0) read (see file header).                         */
/* ------------------------------------------------------------------ */
static void sort_ex(int n, double *s, double *r, double *t, int *ir, int *it)
{
    /* ascending by s; stable insertion sort (ties keep input order) */
    int i, j;
    for (i = 2; i <= n; i++) {
        double ks = s[i], kr = r[i], kt = t[i];
        int ki = ir[i], kj = it[i];
        for (j = i - 1; j >= 1 && s[j] > ks; j--) {
            s[j + 1] = s[j]; r[j + 1] = r[j]; t[j + 1] = t[j];
            ir[j + 1] = ir[j]; it[j + 1] = it[j];
        }
        s[j + 1] = ks; r[j + 1] = kr; t[j + 1] = kt;
        ir[j + 1] = ki; it[j + 1] = kj;
    }
}

static double theta_of_s(double x0, double z0, double sinth02,
                         double costh0, double s)
{
    double th = atan(sqrt(x0 * x0 + sinth02 * s * s) / (z0 + costh0 * s));
    if (th < 0.0) th = th + PICONST;
    return th;
}

int orc_trajectory(orc_traj *t, const orc_grid *g, double rstar,
                   double x0, double z0, double theta0, double rs0_fill)
{
    const double eps = TELESC_EPS, epsplus = 1.0e1 * TELESC_EPS;
    const double pi = PICONST, send = 1.0e30;
    int nr = g->nr, ny = g->nth;
    int iyeq = ny / 2, iyend = ny;
    int nmax = 2 * (nr + 2) + ny + RAYEXPT + 8;
    double *rrgrid = g->rc + 1, *ttgrid = g->tc + 1; /* 1-based views  */
    double pitheta0, costh0, sinth0, costh02, sinth02;
    double *sar1, *sar2, *th_radius, *th_theta, *th_s, *r_radius, *r_theta, *r_s;
    int *iyar, *th_ir, *th_itheta, *r_ir, *r_itheta;
    double ex_s[RAYEXPT + 2], ex_radius[RAYEXPT + 2], ex_theta[RAYEXPT + 2];
    int ex_ir[RAYEXPT + 2], ex_itheta[RAYEXPT + 2];
    int is, iss, iy, ix, isnr, isdblnr, iup, ith_amount, ir_amount, ir_min;
    int isex, nrex, ist, isr, tr_total, rc = 0;
    double bimpact, rmaxt, rmaxr, rmint, rminr, sbeg, sprev;

    sar1 = (double *)calloc(nmax, sizeof(double));
    sar2 = (double *)calloc(nmax, sizeof(double));
    th_radius = (double *)calloc(nmax, sizeof(double));
    th_theta = (double *)calloc(nmax, sizeof(double));
    th_s = (double *)calloc(nmax, sizeof(double));
    r_radius = (double *)calloc(nmax, sizeof(double));
    r_theta = (double *)calloc(nmax, sizeof(double));
    r_s = (double *)calloc(nmax, sizeof(double));
    iyar = (int *)calloc(nmax, sizeof(int));
    th_ir = (int *)calloc(nmax, sizeof(int));
    th_itheta = (int *)calloc(nmax, sizeof(int));
    r_ir = (int *)calloc(nmax, sizeof(int));
    r_itheta = (int *)calloc(nmax, sizeof(int));

    pitheta0 = 0.5 * pi - theta0;
    costh0 = cos(theta0);
    sinth0 = sin(theta0);
    costh02 = costh0 * costh0;
    sinth02 = sinth0 * sinth0;

    /* theta=const cones, :2968-3000 */
    is = 1;
    for (iy = 1; iy <= iyeq; iy++) {
        double theta = TC(g, iy), tn = tan(theta), tanth2 = tn * tn;
        double a = tanth2 * costh02 - sinth02;
        double b = 2.0 * tanth2 * costh0 * z0;
        double c = tanth2 * z0 * z0 - x0 * x0;
        double sd = b * b - 4.0 * a * c;
        if (sd > 0.0) {
            sd = sqrt(sd);
            iyar[is] = iy;
            sar1[is] = (-b - sd) / (2.0 * a);
            sar2[is] = (-b + sd) / (2.0 * a);
            if (sar1[is] > sar2[is]) {
                double d = sar1[is]; sar1[is] = sar2[is]; sar2[is] = d;
            }
            is++;
        }
    }
    isnr = is - 1;
    isdblnr = 0;
    for (is = 1; is <= isnr; is++) {             /* :3008-3014         */
        double pit = 0.5 * pi - TC(g, iyar[is]);
        if (fabs(pit) > fabs(pitheta0)) isdblnr++;
    }
    iup = pitheta0 > 0.0 ? 1 : 0;
#define THR(S) sqrt(x0 * x0 + z0 * z0 + (S) * (S) + 2.0 * z0 * costh0 * (S))
#define TH_LOWQ(I) (iup == 1 ? iyend + 1 - iyar[I] : iyar[I])
#define TH_UPQ(I) (iup == 1 ? iyar[I] : iyend + 1 - iyar[I])
#define PUT(ISS, SV, ITV) do { th_radius[ISS] = THR(SV); th_s[ISS] = (SV); \
        th_itheta[ISS] = (ITV); th_theta[ISS] = TC(g, th_itheta[ISS]); } while (0)
    if (z0 * pitheta0 > 0.0) {                   /* :3024-3108         */
        is = 1 + isdblnr;
        for (iss = 1; iss <= isnr - isdblnr; iss++) {
            PUT(iss, sar1[is], TH_LOWQ(is)); is++;
        }
        is = isnr;
        for (iss = isnr - isdblnr + 1; iss <= 2 * isnr - 2 * isdblnr; iss++) {
            PUT(iss, sar2[is], TH_UPQ(is)); is--;
        }
        if (isdblnr > 0) {
            for (iss = 2 * isnr + 1 - 2 * isdblnr; iss <= 2 * isnr - isdblnr; iss++) {
                PUT(iss, sar1[is], TH_UPQ(is)); is--;
            }
            is++;
            for (iss = 2 * isnr + 1 - isdblnr; iss <= 2 * isnr; iss++) {
                PUT(iss, sar2[is], TH_UPQ(is)); is++;
            }
        }
    } else {                                     /* :3109-3194         */
        is = isdblnr;
        if (isdblnr > 0) {
            for (iss = 1; iss <= isdblnr; iss++) {
                PUT(iss, sar1[is], TH_LOWQ(is)); is--;
            }
            is++;
            for (iss = isdblnr + 1; iss <= 2 * isdblnr; iss++) {
                PUT(iss, sar2[is], TH_LOWQ(is)); is++;
            }
        }
        is = isdblnr + 1;
        for (iss = 2 * isdblnr + 1; iss <= isnr + isdblnr; iss++) {
            PUT(iss, sar1[is], TH_LOWQ(is)); is++;
        }
        is = isnr;
        for (iss = isnr + isdblnr + 1; iss <= 2 * isnr; iss++) {
            PUT(iss, sar2[is], TH_UPQ(is)); is--;
        }
    }
#undef PUT
    ith_amount = 2 * isnr;
    for (is = 1; is <= ith_amount; is++) {        /* :3204-3214         */
        double r = th_radius[is];
        if (r < RC(g, 1)) th_ir[is] = 0;
        else if (r > RC(g, nr)) th_ir[is] = nr;
        else th_ir[is] = orc_hunt(rrgrid, nr, r, 0);
    }

    /* R=const spheres, :3241-3335 */
    bimpact = sqrt(x0 * x0 + z0 * z0 * (1.0 - costh02));
    for (ix = 1; ix <= nr; ix++)
        if (RC(g, ix) > bimpact) break;
    if (ix > nr) { rc = 13; goto done; }
    ir_min = ix;
    if (bimpact <= rstar) ir_min = 0;
    t->b = bimpact;
    ir_amount = 2 * (nr + 1 - ir_min);
    is = 1;
    for (ix = nr; ix >= ir_min; ix--) {
        double r = ix == 0 ? rstar : RC(g, ix);
        double a = 1.0, b = 2.0 * costh0 * z0, c = x0 * x0 + z0 * z0 - r * r;
        double sd = b * b - 4.0 * a * c + eps * b * b, s1, s2;
        if (sd < 0.0) { rc = 13; goto done; }
        sd = sqrt(sd);
        s1 = (-b - sd) / (2.0 * a);
        s2 = (-b + sd) / (2.0 * a);
        r_theta[is] = theta_of_s(x0, z0, sinth02, costh0, s1);
        r_s[is] = s1; r_ir[is] = ix; r_radius[is] = r;
        r_theta[ir_amount + 1 - is] = theta_of_s(x0, z0, sinth02, costh0, s2);
        r_s[ir_amount + 1 - is] = s2;
        r_ir[ir_amount + 1 - is] = ix;
        r_radius[ir_amount + 1 - is] = r;
        is++;
    }
    for (is = 1; is <= ir_amount; is++)
        r_itheta[is] = orc_hunt(ttgrid, ny, r_theta[is], 0);

    tr_total = ir_amount + ith_amount;
    th_s[ith_amount + 1] = 1.0e30;
    r_s[ir_amount + 1] = 1.0e30;
    r_s[0] = rs0_fill;
    rmaxt = RC(g, nr) * (1.0 - epsplus);
    rmaxr = RC(g, nr) * (1.0 + epsplus);
    rmint = RC(g, 1) * (1.0 + epsplus);
    rminr = RC(g, 1) * (1.0 - epsplus);
    sbeg = -1.0e30;
    sprev = -1.0e30;

    /* extra points at the R- and Theta-extrema, :3417-3606 */
    for (isex = 1; isex <= RAYEXPT + 1; isex++) ex_s[isex] = 1.0e30;
    isex = 1;
    if (bimpact > RC(g, 1)) {
        double s2 = 0.0 - z0 * costh0, s3, s1, ds, s0;
        int isrt, irng, iad;
        ex_s[isex] = s2;
        ex_radius[isex] = bimpact;
        ex_theta[isex] = theta_of_s(x0, z0, sinth02, costh0, s2);
        ex_ir[isex] = ir_min - 1;
        ex_itheta[isex] = orc_hunt(ttgrid, ny, ex_theta[isex], 0);
        isex++;
        s3 = s2;
        isrt = orc_hunt(r_s, ir_amount, s2, 0); /* hunt0, nrecip.F:206 */
        for (irng = 1; irng <= 4; irng++) {
            if (irng == 1) { s1 = s3; s2 = r_s[isrt + 1]; }
            else if (irng == 2) { s1 = r_s[isrt]; s2 = s3; }
            else if (irng == 3) { s1 = r_s[isrt + 1]; s2 = r_s[isrt + 2]; }
            else { s1 = r_s[isrt - 1]; s2 = r_s[isrt]; }
            ds = (s2 - s1) / (1.0 + 1.0 * RAYADPT);
            for (iad = 1; iad <= RAYADPT; iad++) {
                double rr;
                s0 = iad * ds + s1;
                ex_s[isex] = s0;
                rr = THR(s0);
                ex_radius[isex] = rr;
                ex_theta[isex] = theta_of_s(x0, z0, sinth02, costh0, s0);
                ex_ir[isex] = orc_hunt(rrgrid, nr, rr, 0);
                ex_itheta[isex] = orc_hunt(ttgrid, ny, ex_theta[isex], 0);
                isex++;
            }
        }
    }
    {
        double s2 = x0 * x0 * costh0 / (z0 * sinth02);
        double rr = THR(s2);
        if (rr > RC(g, 1) && rr < RC(g, nr)) {
            double s3, s1, ds, s0;
            int isrt, irng, iad, ixh;
            ex_s[isex] = s2;
            ex_radius[isex] = rr;
            ex_theta[isex] = theta_of_s(x0, z0, sinth02, costh0, s2);
            ixh = orc_hunt(rrgrid, nr, rr, 0);
            if (ixh == 0 || ixh == nr) { rc = 192; goto done; }
            ex_ir[isex] = ixh;
            ex_itheta[isex] = orc_hunt(ttgrid, ny, ex_theta[isex], 0);
            isex++;
            s3 = s2;
            isrt = orc_hunt(th_s, ith_amount, s3, 0);
            if (!(isrt - RAYRNPT + 1 < 1 || isrt + RAYRNPT > ith_amount)) {
                for (irng = 1; irng <= 4; irng++) {
                    if (irng == 1) { s1 = s3; s2 = th_s[isrt + 1]; }
                    else if (irng == 2) { s1 = th_s[isrt]; s2 = s3; }
                    else if (irng == 3) { s1 = th_s[isrt + 1]; s2 = th_s[isrt + 2]; }
                    else { s1 = th_s[isrt - 1]; s2 = th_s[isrt]; }
                    if (s2 == s1) { rc = 987; goto done; }
                    if (s2 < s1) { rc = 988; goto done; }
                    ds = (s2 - s1) / (1.0 + 1.0 * RAYADPT);
                    for (iad = 1; iad <= RAYADPT; iad++) {
                        double r2;
                        s0 = iad * ds + s1;
                        ex_s[isex] = s0;
                        r2 = THR(s0);
                        ex_radius[isex] = r2;
                        ex_theta[isex] = theta_of_s(x0, z0, sinth02, costh0, s0);
                        ixh = orc_hunt(rrgrid, nr, r2, 0);
                        if (ixh == 0 || ixh == nr) continue;  /* :3580 */
                        ex_ir[isex] = ixh;
                        ex_itheta[isex] = orc_hunt(ttgrid, ny, ex_theta[isex], 0);
                        isex++;
                    }
                }
            }
        }
    }
#undef THR
    nrex = isex - 1;
    sort_ex(nrex, ex_s, ex_radius, ex_theta, ex_ir, ex_itheta);
    isex = 1;

    /* merge, :3610-3677 */
    t->n = 0;
    t->s = (double *)malloc(sizeof(double) * (tr_total + RAYEXPT + 2));
    t->r = (double *)malloc(sizeof(double) * (tr_total + RAYEXPT + 2));
    t->theta = (double *)malloc(sizeof(double) * (tr_total + RAYEXPT + 2));
    t->mu = (double *)malloc(sizeof(double) * (tr_total + RAYEXPT + 2));
    t->phi = (double *)malloc(sizeof(double) * (tr_total + RAYEXPT + 2));
    t->icross = (int *)malloc(sizeof(int) * (tr_total + RAYEXPT + 2));
    t->ir = (int *)malloc(sizeof(int) * (tr_total + RAYEXPT + 2));
    t->it = (int *)malloc(sizeof(int) * (tr_total + RAYEXPT + 2));
    ist = isr = 1;
    is = 0;
#define ADD(IC, R, TH, IR, IT, S) do { t->icross[is] = (IC); t->r[is] = (R); \
        t->theta[is] = (TH); t->ir[is] = (IR); t->it[is] = (IT); t->s[is] = (S); \
        sprev = (S); is++; } while (0)
    for (iss = 1; iss <= tr_total; iss++) {
        for (;;) {
            double m = th_s[ist] < r_s[isr] ? th_s[ist] : r_s[isr];
            if (!(ex_s[isex] < m)) break;
            if (ex_radius[isex] <= rmaxt && ex_radius[isex] >= rmint &&
                (ex_s[isex] - sprev) > eps * ex_radius[isex] &&
                ex_s[isex] >= sbeg && ex_s[isex] <= send)
                ADD(3, ex_radius[isex], ex_theta[isex], ex_ir[isex],
                    ex_itheta[isex], ex_s[isex]);
            isex++;
        }
        if (isex > RAYEXPT) { rc = 83459; goto done; }
        if (th_s[ist] < r_s[isr]) {
            if (th_radius[ist] <= rmaxt && th_radius[ist] >= rmint &&
                (th_s[ist] - sprev) > eps * th_radius[ist] &&
                th_s[ist] >= sbeg && th_s[ist] <= send)
                ADD(2, th_radius[ist], th_theta[ist], th_ir[ist],
                    th_itheta[ist], th_s[ist]);
            ist++;
        } else {
            if (r_radius[isr] <= rmaxr && r_radius[isr] >= rminr &&
                (r_s[isr] - sprev) > eps * r_radius[isr] &&
                r_s[isr] >= sbeg && r_s[isr] <= send)
                ADD(1, r_radius[isr], r_theta[isr], r_ir[isr],
                    r_itheta[isr], r_s[isr]);
            isr++;
        }
    }
#undef ADD
    t->n = is;

    /* local direction (mu, phi), :3682-3718 */
    {
        double znew = z0 * sinth02, bnew = sqrt(x0 * x0 + z0 * z0 * sinth02);
        for (is = 0; is < t->n; is++) {
            double snew = t->s[is] + z0 * costh0, q, dummy, sinphi;
            t->mu[is] = snew / sqrt(bnew * bnew + snew * snew);
            q = znew + snew * costh0;
            dummy = bnew * sqrt(bnew * bnew + snew * snew - q * q);
            if (dummy > 0.0)
                sinphi = (bnew * bnew * costh0 - znew * snew) / dummy;
            else
                sinphi = epsplus;
            if (x0 < 0.0) sinphi = asin(sinphi);
            else sinphi = pi - asin(sinphi);
            while (sinphi < 0.0) sinphi = sinphi + 2.0 * pi;
            while (sinphi >= 2.0 * pi) sinphi = sinphi - 2.0 * pi;
            t->phi[is] = sinphi;
        }
    }
done:
    free(sar1); free(sar2); free(th_radius); free(th_theta); free(th_s);
    free(r_radius); free(r_theta); free(r_s); free(iyar); free(th_ir);
    free(th_itheta); free(r_ir); free(r_itheta);
    return rc;
}

void orc_traj_free(orc_traj *t)
{
    free(t->s); free(t->r); free(t->theta); free(t->mu); free(t->phi);
    free(t->icross); free(t->ir); free(t->it);
    memset(t, 0, sizeof(*t));
}

/* ------------------------------------------------------------------ */
/* Profile, quadrature, element integrators                            */
/* ------------------------------------------------------------------ */
double orc_voigt(double nu00, double aks, double dnu)   /* line.F:2280 */
{
    double nu0 = fabs(nu00);
    double aa = 3.33567e-6 * nu0 * aks;
    double norm = 0.56419583546 / aa;
    double u = dnu / aa;
    return norm * exp(-(u * u));
}

double orc_qdr_src_2(double I, double js1, double alp1, double js2,
                     double alp2, double ds, long *overshoot) /* transfer.F:1498 */
{
    double dtau1 = 0.5 * (alp1 + alp2) * ds;
    double theomax = 0.5 * (js1 + js2) * ds;
    double a, b, xp, src1, src2, q;
    if (dtau1 > 1.0e-6) {
        double e0, e1;
        xp = exp(-dtau1);
        e0 = 1.0 - xp;
        e1 = dtau1 - e0;
        b = e1 / dtau1;
        a = e0 - b;
    } else {
        a = 0.5 * dtau1;
        b = 0.5 * dtau1;
        xp = 1.0 - dtau1;
    }
    if (alp1 > 0.0) src1 = js1 / alp1;
    else if (alp2 > 0.0) src1 = js2 / alp2;
    else src1 = 0.0;
    if (alp2 > 0.0) src2 = js2 / alp2;
    else if (alp1 > 0.0) src2 = js1 / alp1;
    else src2 = 0.0;
    /* the 1e-9 literal is single precision (transfer.F:1542) */
    if (dtau1 > (double)1e-9f) q = a * src1 + b * src2;
    else q = 0.5 * (js1 + js2) * ds;
    if (q > 2.0 * theomax && overshoot) (*overshoot)++;
    q = q < theomax ? q : theomax;
    return I * xp + q;
}

typedef struct {
    double nu0, A, Bud, Bdu, dnu_ch, u_ch;
    double tau, emis;
    long overshoot;
} line_ctx;

typedef struct {               /* line carry across elements            */
    double phiprof0, srcl0, alpl0;
    int init;
} carry_t;

/* integrate_element_linedust, line.F:4515-4624 */
static void integ_elem(const line_ctx *L, double *I, double ds,
                       double srcd0, double srcd1, double alpd0, double alpd1,
                       double lw0, double lw1, double dvmu0, double dvmu1,
                       double nup0, double nup1, double ndown0, double ndown1,
                       carry_t *c, double *tau, double *emis, long *ovs)
{
    const double hk = 5.27296241956e-28;
    double lwav = 0.5 * (lw0 + lw1);
    double dnu0 = L->dnu_ch - L->nu0 * dvmu0;
    double dnu1 = L->dnu_ch - L->nu0 * dvmu1;
    double phi1, srcl1, alpl1, src0, src1, alp0, alp1;
    if (c->init == 1) c->phiprof0 = orc_voigt(L->nu0, lwav, dnu0);
    phi1 = orc_voigt(L->nu0, lwav, dnu1);
    if (c->init == 1) c->srcl0 = hk * L->nu0 * nup0 * L->A * c->phiprof0;
    srcl1 = hk * L->nu0 * nup1 * L->A * phi1;
    if (c->init == 1)
        c->alpl0 = hk * L->nu0 * c->phiprof0 * (ndown0 * L->Bdu - nup0 * L->Bud);
    alpl1 = hk * L->nu0 * phi1 * (ndown1 * L->Bdu - nup1 * L->Bud);
    src0 = srcd0 + c->srcl0;
    src1 = srcd1 + srcl1;
    alp0 = alpd0 + c->alpl0;
    alp1 = alpd1 + alpl1;
    *I = orc_qdr_src_2(*I, src0, alp0, src1, alp1, ds, ovs);
    *tau = *tau + 0.5 * (alp0 + alp1) * ds;
    *emis = *emis + 0.5 * (src0 + src1) * ds;
    c->phiprof0 = phi1;
    c->srcl0 = srcl1;
    c->alpl0 = alpl1;
    c->init = 0;
}

/* clever_integrate_element_linedust, line.F:4636-4848 (SUBGRID on) */
static void clever_elem(const line_ctx *L, double *I, double ds,
                        double srcd0, double srcd1, double alpd0, double alpd1,
                        double lw0, double lw1, double dvmu0, double dvmu1,
                        double nup0, double nup1, double ndown0, double ndown1,
                        carry_t *c, double *tau, double *emis, long *ovs)
{
    const double crit = 3.0;
    enum { NMAX = 31 };
    double lw = 0.5 * (lw0 + lw1);
    double x = fabs((dvmu1 - dvmu0) / (lw / 2.99792458e5));
    if (2.0 * crit * x > 1.0) {
        double s_c = ds * (L->u_ch - dvmu0) / (dvmu1 - dvmu0);
        double dls = ds / x;
        double sright = s_c + crit * dls, sleft = s_c - crit * dls;
        if (sright > 0.0 && sleft < ds) {
            double ls[NMAX + 3], lnu[NMAX + 3], lnd[NMAX + 3], ldv[NMAX + 3];
            double lsd[NMAX + 3], lad[NMAX + 3], lg_ds;
            int n = 1, i;
            lg_ds = (sright - sleft) / (NMAX - 1.0);
            ls[1] = 0.0; lnu[1] = nup0; lnd[1] = ndown0; ldv[1] = dvmu0;
            lsd[1] = srcd0; lad[1] = alpd0;
            for (i = 1; i <= NMAX; i++) {
                double s = sleft + (i - 1) * lg_ds;
                if (s > 0.0 && s < ds) {
                    double e = s / ds, ep = 1.0 - e;
                    n++;
                    ls[n] = s;
                    lnu[n] = ep * nup0 + e * nup1;
                    lnd[n] = ep * ndown0 + e * ndown1;
                    ldv[n] = ep * dvmu0 + e * dvmu1;
                    lsd[n] = ep * srcd0 + e * srcd1;
                    lad[n] = ep * alpd0 + e * alpd1;
                }
            }
            n++;
            ls[n] = ds; lnu[n] = nup1; lnd[n] = ndown1; ldv[n] = dvmu1;
            lsd[n] = srcd1; lad[n] = alpd1;
            for (i = 2; i <= n; i++)
                integ_elem(L, I, ls[i] - ls[i - 1], lsd[i - 1], lsd[i],
                           lad[i - 1], lad[i], lw, lw, ldv[i - 1], ldv[i],
                           lnu[i - 1], lnu[i], lnd[i - 1], lnd[i], c, tau, emis, ovs);
            return;
        }
    }
    integ_elem(L, I, ds, srcd0, srcd1, alpd0, alpd1, lw0, lw1, dvmu0, dvmu1,
               nup0, nup1, ndown0, ndown1, c, tau, emis, ovs);
}

/* ------------------------------------------------------------------ */
/* Node values: get_line_dust_values (line.F:3965-4217) + omega_dot_v  */
/* (line.F:2650-2662).                                                 */
/* ------------------------------------------------------------------ */
typedef struct { double src, alp, dvmu, lw, nup, ndown; } node_t;

static int node_values(const orc_model *m, const orc_grid *g, int il,
                       int icr, int ir, int it, double dr, double dt,
                       double mu, double phi, node_t *o)
{
    int nh = m->nth_half, nl = m->nlines, nlev = m->nlev, ny = g->nth;
    int r0, r1, t0, t1, k;
    int lu = m->lev_up[il], ld = m->lev_low[il];
    double molpg = 1.0 / (m->umass_av * 1.6726e-24);
    double v[3];
    if (dr < 0.0 || dr > 1.0) return 6024;
    if (dt < 0.0 || dt > 1.0) return 6023;
    if (dr > 0.0) { r0 = ir; r1 = ir + 1; if (r1 > m->nr) r1 = m->nr; }
    else { r0 = ir; r1 = ir - 1; if (r1 < 1) r1 = 1; }
    if (dt > 0.0) { t0 = it; t1 = it + 1; } else { t0 = it; t1 = it - 1; }
    t0 = ridx_it(t0, ny);
    t1 = ridx_it(t1, ny);
#define CELL(IR, IT) (((IR) - 1) * nh + ((IT) - 1))
#define JD(IR, IT) m->jd[CELL(IR, IT) * nl + il]
#define AD(IR, IT) m->ad[CELL(IR, IT) * nl + il]
#define LW(IR, IT) m->lw[CELL(IR, IT)]
#define NN(L, IR, IT) (m->pops[CELL(IR, IT) * nlev + (L)] * m->abund[CELL(IR, IT)] \
                       * m->rho[CELL(IR, IT)] * molpg)
#define VV(K, IR, IT) m->vel[CELL(IR, IT) * 3 + (K)]
    if (icr == 1) {
        double w0 = 1.0 - dt, w1 = dt;
        o->src = w0 * JD(r0, t0) + w1 * JD(r0, t1);
        o->alp = w0 * AD(r0, t0) + w1 * AD(r0, t1);
        o->lw = w0 * LW(r0, t0) + w1 * LW(r0, t1);
        o->nup = w0 * NN(lu, r0, t0) + w1 * NN(lu, r0, t1);
        o->ndown = w0 * NN(ld, r0, t0) + w1 * NN(ld, r0, t1);
        for (k = 0; k < 3; k++) v[k] = w0 * VV(k, r0, t0) + w1 * VV(k, r0, t1);
    } else if (icr == 2) {
        double w0 = 1.0 - dr, w1 = dr;
        o->src = w0 * JD(r0, t0) + w1 * JD(r1, t0);
        o->alp = w0 * AD(r0, t0) + w1 * AD(r1, t0);
        o->lw = w0 * LW(r0, t0) + w1 * LW(r1, t0);
        o->nup = w0 * NN(lu, r0, t0) + w1 * NN(lu, r1, t0);
        o->ndown = w0 * NN(ld, r0, t0) + w1 * NN(ld, r1, t0);
        for (k = 0; k < 3; k++) v[k] = w0 * VV(k, r0, t0) + w1 * VV(k, r1, t0);
    } else {
#define BIL(F) ((1.0 - dr) * ((1.0 - dt) * F(r0, t0) + dt * F(r0, t1)) + \
                dr * ((1.0 - dt) * F(r1, t0) + dt * F(r1, t1)))
#define NU_(IR, IT) NN(lu, IR, IT)
#define ND_(IR, IT) NN(ld, IR, IT)
#define V0_(IR, IT) VV(0, IR, IT)
#define V1_(IR, IT) VV(1, IR, IT)
#define V2_(IR, IT) VV(2, IR, IT)
        o->src = BIL(JD);
        o->alp = BIL(AD);
        o->lw = BIL(LW);
        o->nup = BIL(NU_);
        o->ndown = BIL(ND_);
        v[0] = BIL(V0_);
        v[1] = BIL(V1_);
        v[2] = BIL(V2_);
#undef BIL
    }
    if (mu > 1.0) return 393;
    o->dvmu = 3.335668e-11 * (mu * v[0] + sqrt(1.0 - mu * mu) *
                               (v[1] * sin(phi) + v[2] * cos(phi)));
    return 0;
#undef CELL
#undef JD
#undef AD
#undef LW
#undef NN
#undef VV
}

/* per-ray node cache: node values do not depend on the channel */
typedef struct {
    int n;
    double *ds;        /* ds[k] = s_k - s_{k-1}, k>=1                    */
    int *reset;        /* hole chord: ds := 0 and init := 1              */
    int *blend;        /* star blend happens at node k (central beam)   */
    int *star;         /* I := I_star at node k (rbeam0 = 0, b <= R*)    */
    double starfract;  /* (rstar/rbeam0)^2                              */
    node_t *nv;
} raycache_t;

/* charintline, telescope.F:3889-4312 (rbeam0 > 0 only for the centre) */
static double charint(const raycache_t *rc_, line_ctx *L,
                      double istar_ch, double I0, long *ovs, int *maser)
{
    double I = I0;
    carry_t c = {0.0, 0.0, 0.0, 1};
    int k;
    double tau = 0.0, emis = 0.0;
    const node_t *a = &rc_->nv[0];
    for (k = 1; k < rc_->n; k++) {
        const node_t *b = &rc_->nv[k];
        double ds = rc_->ds[k];
        if (rc_->reset[k]) { ds = 0.0; c.init = 1; }
        if (rc_->blend[k]) {            /* telescope.F:4187-4193 */
            I = (1.0 - rc_->starfract) * I + rc_->starfract * istar_ch;
            c.init = 1;
        }
        if (rc_->star[k]) {             /* telescope.F:4194-4208 */
            I = istar_ch;
            c.init = 1;
        }
        clever_elem(L, &I, ds, a->src, b->src, a->alp, b->alp, a->lw, b->lw,
                    a->dvmu, b->dvmu, a->nup, b->nup, a->ndown, b->ndown,
                    &c, &tau, &emis, ovs);
        if (c.alpl0 * ds < -0.01) *maser = 1;
        a = b;
    }
    L->tau = tau;
    L->emis = emis;
    return I;
}

/* Build the channel-independent node list of one ray for one line,
 * with the hole / star-blend decisions of charintline
 * (telescope.F:4050-4208).  rbeam0 > 0 marks the central beam. */
static int build_cache(const orc_model *m, const orc_grid *g,
                       const orc_traj *t, int il, double rbeam0,
                       raycache_t *rc_)
{
    int k, rc = 0, istar_done = 0;
    int nr = g->nr;
    rc_->n = t->n;
    rc_->ds = (double *)calloc(t->n + 1, sizeof(double));
    rc_->reset = (int *)calloc(t->n + 1, sizeof(int));
    rc_->blend = (int *)calloc(t->n + 1, sizeof(int));
    rc_->star = (int *)calloc(t->n + 1, sizeof(int));
    rc_->nv = (node_t *)calloc(t->n + 1, sizeof(node_t));
    (void)nr;
    for (k = 0; k < t->n; k++) {
        int ir = t->ir[k], it = t->it[k];
        double dr, dt;
        if (ir == 0) return 7454;
        dr = (t->r[k] - RC(g, ir)) / (RC(g, ir + 1) - RC(g, ir));
        dt = (t->theta[k] - TC(g, it)) / (TC(g, it + 1) - TC(g, it));
        rc = node_values(m, g, il, t->icross[k], ir, it, dr, dt, t->mu[k],
                         t->phi[k], &rc_->nv[k]);
        if (rc) return rc;
        if (k == 0) continue;
        rc_->ds[k] = t->s[k] - t->s[k - 1];
        if (rc_->ds[k] < 0.0) return 749;
        if (ir == 1 && t->ir[k - 1] == 1 && t->icross[k] == 1 && t->icross[k - 1] == 1)
            rc_->reset[k] = 1;
        if (ir <= 1 && istar_done == 0 && rbeam0 > 0.0) {
            if (rbeam0 < m->rstar) return 124;
            rc_->blend[k] = 1;
            istar_done = 1;
        }
        /* a ray of the rectangular imager (rbeam0 = 0) whose impact
           parameter is within the star takes the stellar intensity at
           every node following a node of the innermost radial cell,
           itself in that cell (telescope.F:4194-4208) */
        if (ir == 1 && t->ir[k - 1] == 1 && rbeam0 == 0.0 && t->b <= m->rstar)
            rc_->star[k] = 1;
    }
    if (rbeam0 > 0.0) {
        double q = m->rstar / rbeam0;
        rc_->starfract = q * q;
    }
    return 0;
}

static void free_cache(raycache_t *c)
{
    free(c->ds); free(c->reset); free(c->blend); free(c->star); free(c->nv);
}

/* ------------------------------------------------------------------ */
/* calc_line_spectrum + make_image_circular + calc_freq_flux_observer  */
/* (telescope.F:1465-1654, 339-631, 1320-1437)                         */
/* ------------------------------------------------------------------ */
int orc_passband_nfr(double vmax, double dv, double *vmax_out)
{
    int nfr = (int)(2.0 * vmax / dv + 1);   /* main.F:209 (truncation) */
    *vmax_out = 0.5 * (nfr - 1) * dv;        /* main.F:210             */
    return nfr;
}

static double star_at(const orc_model *m, double freq)
{
    /* line_init_passband_starbc, line.F:3821-3843 */
    const double *c = m->cont_nu - 1;          /* 1-based view          */
    const double *s = m->cont_istar - 1;
    int n = m->ncont, i;
    double wgt;
    if (n < 2) return 0.0;
    i = orc_hunt(c, n, freq, 0);
    if (i == 0 || i == n) return 0.0;
    wgt = (freq - c[i]) / (c[i + 1] - c[i]);
    return (1.0 - wgt) * s[i] + wgt * s[i + 1];
}

int orc_line_spectrum(const orc_model *m, const orc_grid *g,
                      const orc_rays *rays, int il, double width,
                      int nfr, int skip_mode, int nthreads, int ray_stride,
                      double *flux, double *image, int *cmask,
                      float *minvel, float *maxvel, double *dnu_out,
                      int *maser_out, long *overshoot_out, long *nint_out)
{
    double nu0 = fabs(m->nu0[il]);
    double passb, nu1, dnu, aksmax_c, *dnu_ch, *velo, *istar, *img;
    double Bud, Bdu, *I_bc;
    int nrays = rays->nrays, k, err = 0, maser = 0;
    long ovs = 0, nint = 0;
    const double D = 1.0 * 3.08572e18;         /* main.F:206,213         */

    if (nfr < 2) return 13;
    /* line_setup_passband, line.F:462-469 */
    passb = 3.33567e-6 * nu0 * width;
    nu1 = 0.0 - passb;
    dnu = 2.0 * passb / (nfr - 1.0);
    dnu_ch = (double *)malloc(sizeof(double) * nfr);
    velo = (double *)malloc(sizeof(double) * nfr);
    istar = (double *)malloc(sizeof(double) * nfr);
    for (k = 0; k < nfr; k++) {
        dnu_ch[k] = nu1 + k * dnu;
        velo[k] = dnu_ch[k] / m->nu0[il];          /* telescope.F:392   */
        istar[k] = star_at(m, nu0 + dnu_ch[k]);
        if (dnu_out) dnu_out[k] = dnu_ch[k];
    }
    /* prepare_lines, line.F:1778-1787 */
    Bud = 6.78171833781e46 * m->A[il] / (m->nu0[il] * m->nu0[il] * m->nu0[il]);
    Bdu = Bud * (m->gu[il] / m->gl[il]);
    aksmax_c = m->aksmax / 2.99792458e5;
    I_bc = (double *)malloc(sizeof(double) * nfr);
    orc_outer_bc(m, il, nfr, dnu_ch, I_bc);
    img = image ? image : (double *)malloc(sizeof(double) * (size_t)nrays * nfr);

#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+:ovs, nint) reduction(|:maser)
#endif
    for (int iray = 0; iray < nrays; iray++) {
        orc_traj t;
        if (ray_stride > 1 && iray % ray_stride != 0) continue;
        raycache_t cache;
        line_ctx L;
        int e, ch, mz = 0;
        float mn = 1.0f, mx = -1.0f;
        double cont = 0.0, rbeam0 = iray == 0 ? rays->ring_ri[1] : -1.0;
        double *row = img + (size_t)iray * nfr;
        memset(&t, 0, sizeof t);
        memset(&cache, 0, sizeof cache);
        e = orc_trajectory(&t, g, m->rstar, rays->x0[iray], rays->z0[iray],
                           rays->theta0, 0.0);
        if (!e) e = build_cache(m, g, &t, il, rbeam0, &cache);
        if (e) {
#ifdef _OPENMP
#pragma omp critical
#endif
            err = e;
            orc_traj_free(&t);
            free_cache(&cache);
            continue;
        }
        L.nu0 = m->nu0[il]; L.A = m->A[il]; L.Bud = Bud; L.Bdu = Bdu;
        /* NONREDUNDANT window (telescope.F:4265-4270): min/max of the
           previous node's projected velocity where N_up+N_low > LEVTHRES,
           kept in REAL*4; identical for every channel, so it is final
           after the first integrated channel. */
        for (k = 0; k + 1 < cache.n; k++) {
            const node_t *p = &cache.nv[k];
            if (p->dvmu < mn && p->nup + p->ndown > LEVTHRES) mn = (float)p->dvmu;
            if (p->dvmu > mx && p->nup + p->ndown > LEVTHRES) mx = (float)p->dvmu;
        }
        for (ch = 0; ch < nfr; ch++) {
            int doit;
            L.dnu_ch = dnu_ch[ch];
            L.u_ch = velo[ch];
            if (iray == 0 || skip_mode == 1 || ch == 0) doit = 1;
            else doit = (velo[ch] <= (double)mx + 2.0 * aksmax_c &&
                         velo[ch] >= (double)mn - 2.0 * aksmax_c);
            if (cmask) cmask[(size_t)iray * nfr + ch] = 0;
            if (doit) {
                row[ch] = charint(&cache, &L, istar[ch], I_bc[ch], &ovs, &mz);
                nint++;
                if (cmask) cmask[(size_t)iray * nfr + ch] = 1;
                if (ch == 0 && iray > 0 && skip_mode == 0 &&
                    (velo[0] > (double)mx + 2.0 * aksmax_c ||
                     velo[0] < (double)mn - 2.0 * aksmax_c))
                    cont = row[0];
            } else if (cont != 0.0) {
                row[ch] = cont;
            } else {
                row[ch] = charint(&cache, &L, istar[ch], I_bc[ch], &ovs, &mz);
                nint++;
                cont = row[ch];
            }
        }
        if (mz) maser = 1;
        if (minvel) minvel[iray] = mn;
        if (maxvel) maxvel[iray] = mx;
        orc_traj_free(&t);
        free_cache(&cache);
    }
    if (!err && ray_stride <= 1) {
        /* calc_freq_flux_observer, telescope.F:1386-1435 */
        const double *ri = rays->ring_ri;
        int nphi = rays->nphi, nrings = rays->nrings;
        for (k = 0; k < nfr; k++) {
            double slum = 0.0, surf, dslum;
            int ir, ip;
            surf = 3.14159265359 * (ri[1] * ri[1]);
            slum = slum + surf * img[k];
            for (ir = 1; ir <= nrings; ir++) {
                surf = 3.14159265359 * (ri[ir + 1] * ri[ir + 1] - ri[ir] * ri[ir]);
                dslum = 0.0;
                for (ip = 0; ip < nphi; ip++)
                    dslum = dslum + img[(size_t)(1 + (ir - 1) * nphi + ip) * nfr + k];
                dslum = dslum / (1.0 * nphi);
                dslum = dslum * surf;
                slum = slum + dslum;
            }
            flux[k] = slum / (D * D);
        }
    }
    if (maser_out) *maser_out = maser;
    if (overshoot_out) *overshoot_out = ovs;
    if (nint_out) *nint_out = nint;
    if (!image) free(img);
    free(dnu_ch); free(velo); free(istar); free(I_bc);
    return err;
}

void orc_outer_bc(const orc_model *m, int il, int nfr, const double *dnu_ch, double *I_bc)
{
    const double nu0 = fabs(m->nu0[il]);
    int k;
    for (k = 0; k < nfr; k++) I_bc[k] = 0.0;              /* type 0: black */
    if (m->outer_bc == 2) {                                /* telescope.F:3996-4005 */
        const double f = m->nu0[il];
        const double v = 1.47455253991e-47 * (f * f * f) / (exp(4.7991598e-11 * f / 2.728) - 1.0);
        for (k = 0; k < nfr; k++) I_bc[k] = v;
    } else if (m->outer_bc == 3 && m->cont_isrf) {
        /* line_init_passband_outerbc, line.F:3878-3900: the channel
           frequency freq_nu = nu0 + dnu (line.F:466-467, 518) on the
           continuum grid.  The upper-end test compares the interval with
           freq_nr, which line_setup_passband has just set to the
           passband's channel count (line.F:505); an interval that ends at
           the last continuum point reads the point after it, which is 0
           in the reference's (static, zero-filled) arrays. */
        const double *c = m->cont_nu - 1, *f = m->cont_isrf - 1;
        const int n = m->ncont;
        for (k = 0; k < nfr; k++) {
            const double freq = nu0 + dnu_ch[k];
            const int i = orc_hunt(c, n, freq, 0);
            double wgt, c1, f1;
            if (i == 0 || i == nfr) continue;
            c1 = i < n ? c[i + 1] : 0.0;
            f1 = i < n ? f[i + 1] : 0.0;
            wgt = (freq - c[i]) / (c1 - c[i]);
            I_bc[k] = (1.0 - wgt) * f[i] + wgt * f1;
        }
    }
}

/* ------------------------------------------------------------------ */
/* R13 table builders                                                  */
/* ------------------------------------------------------------------ */
double orc_bplanck(double temp, double nu)             /* setup.F:937 */
{
    if (temp == 0.0) return 0.0;
    return 1.47455e-47 * nu * nu * nu / (exp(4.7989e-11 * nu / temp) - 1.0) + 1.0e-290;
}

void orc_line_consts(int nlev, const double *E_cm, int nlines,
                     const int *lev_up, const int *lev_low, double *nu0)
{
    /* read_species_lambda, line.F:1902-1903, 1980-1983 */
    int i;
    (void)nlev;
    for (i = 0; i < nlines; i++) {
        double eu = 1.986468498e-16 * E_cm[lev_up[i]];
        double el = 1.986468498e-16 * E_cm[lev_low[i]];
        nu0[i] = 1.509160e26 * (eu - el);
    }
}

int orc_line_dust(int nr, int nth, int nspec, const double *rho_d,
                  const double *t_d, int nf, const double *cont_nu,
                  const double *kabs, const double *kscat, const double *scat,
                  int nlines, const double *nu0, double *jd, double *ad)
{
    /* line_dust_compute_src_alp, line.F:3560-3607 (grey per line) */
    const double *c = cont_nu - 1;
    int cell, il, isp;
    for (cell = 0; cell < nr * nth; cell++) {
        for (il = 0; il < nlines; il++) {
            double freq = nu0[il], wgt, src = 0.0, alp = 0.0;
            int i = orc_hunt(c, nf, freq, 0);
            if (i == 0 || i == nf) {
                jd[(size_t)cell * nlines + il] = 0.0;
                ad[(size_t)cell * nlines + il] = 0.0;
                continue;
            }
            wgt = (freq - c[i]) / (c[i + 1] - c[i]);
            for (isp = 0; isp < nspec; isp++) {
                const double *ka = kabs + (size_t)isp * nf - 1;
                double kw = wgt * ka[i + 1] + (1.0 - wgt) * ka[i];
                double rd = rho_d[(size_t)isp * nr * nth + cell];
                src = src + rd * kw * orc_bplanck(t_d[(size_t)isp * nr * nth + cell], freq);
                alp = alp + rd * kw;
            }
            /* scattering/PAH source, line.F:3593-3594: scati_src from
               scatsource.dat + quantsource.dat (source.F:494-546), 0 when
               neither file exists (COMMON storage starts at 0) */
            {
                const double s1 = scat ? scat[(size_t)i * nr * nth + cell] : 0.0;
                const double s0 = scat ? scat[(size_t)(i - 1) * nr * nth + cell] : 0.0;
                src = src + wgt * s1 + (1.0 - wgt) * s0;
            }
            for (isp = 0; isp < nspec; isp++) {
                const double *ks = kscat + (size_t)isp * nf - 1;
                double kw = wgt * ks[i + 1] + (1.0 - wgt) * ks[i];
                alp = alp + rho_d[(size_t)isp * nr * nth + cell] * kw;
            }
            jd[(size_t)cell * nlines + il] = src;
            ad[(size_t)cell * nlines + il] = alp;
        }
    }
    return 0;
}

void orc_star_intensity(int nf, const double *fnu, double rstar, double *is)
{
    int i;   /* read_starspec, star.F:508-525 */
    for (i = 0; i < nf; i++) is[i] = 3.0308410e36 * fnu[i] / (rstar * rstar);
}

double orc_aksmax(int ncell, const double *lw, double fix_lw)
{
    double mx = 0.0;  /* line_velo_prep_profiles_abun, line.F:3020-3033 */
    int i;
    for (i = 0; i < ncell; i++) {
        double a = lw[i];
        if (a <= 0.0) a = fix_lw;
        if (a > mx) mx = a;
    }
    return mx;
}

/* Test hook: the channel-independent node values of one ray for one line
 * (get_line_dust_values at every trajectory node) with the hole-reset and
 * star-blend flags, as charintline consumes them.  Arrays of size >= the
 * node count; returns the count, or minus a stop code. */
int orc_ray_nodes(const orc_model *m, const orc_grid *g, const orc_rays *rays, int il,
                  int iray, int cap, double *ds, double *u, double *jd, double *ad,
                  double *lw, double *nu, double *nl, int *flags)
{
    orc_traj t;
    raycache_t c;
    int e, k, n;
    double rbeam0 = iray == 0 ? rays->ring_ri[1] : -1.0;
    memset(&t, 0, sizeof t);
    memset(&c, 0, sizeof c);
    e = orc_trajectory(&t, g, m->rstar, rays->x0[iray], rays->z0[iray], rays->theta0, 0.0);
    if (!e) e = build_cache(m, g, &t, il, rbeam0, &c);
    n = t.n;
    if (!e && n > cap) e = 13;
    if (!e) {
        for (k = 0; k < n; k++) {
            ds[k] = c.ds[k]; u[k] = c.nv[k].dvmu; jd[k] = c.nv[k].src; ad[k] = c.nv[k].alp;
            lw[k] = c.nv[k].lw; nu[k] = c.nv[k].nup; nl[k] = c.nv[k].ndown;
            flags[k] = (c.reset[k] ? 1 : 0) | (c.blend[k] ? 2 : 0);
        }
        if (iray == 0) ds[0] = c.starfract;
    }
    orc_traj_free(&t);
    free_cache(&c);
    return e ? -e : n;
}

/* ------------------------------------------------------------------ */
/* Rectangular images: setup_rays_rectang (telescope.F:2229-2468) and   */
/* make_freq_image_rectang (telescope.F:2061-2183) as                   */
/* calc_write_line_posvel (telescope.F:1828-2031) uses them.            */
/* ------------------------------------------------------------------ */
int orc_setup_rays_rect(orc_rays *rays, double *b, double anginf, int nx, int ny,
                        double size_x, double size_y, double phioff, double xoff,
                        double yoff, int addstar, double rstar, double *spx_out,
                        double *spy_out, int *starunres)
{
    double szx, szy, spx, spy, theta0, sinth0;
    int ix, iy, ir = 0, nxh, nyh, n;
    if (nx % 2 || ny % 2 || nx < 2 || ny < 2) return 13;          /* :2262-2269 */
    szx = size_x * 0.5;                                            /* :209-212  */
    szy = size_y * 0.5;
    spx = szx / (nx * 0.5);
    spy = szy / (ny * 0.5);
    if (spx <= 0.0 || spy <= 0.0) return 13;
    if (anginf < 1.0e-1) anginf = (double)0.1f;     /* :2294-2303, a REAL literal */
    theta0 = anginf + 1.0e-4;
    sinth0 = sin(theta0);
    n = nx * ny + (addstar > 0 ? 1 : 0);
    memset(rays, 0, sizeof(*rays));
    rays->nrays = n;
    rays->theta0 = theta0;
    rays->x0 = (double *)calloc(n, sizeof(double));
    rays->z0 = (double *)calloc(n, sizeof(double));
    nxh = nx / 2;
    nyh = ny / 2;
    for (ix = 1; ix <= nx; ix++) {
        for (iy = 1; iy <= ny; iy++) {
            double x_c = (ix - nxh - 0.5) * spx - xoff;
            double y_c = (iy - nyh - 0.5) * spy - yoff;
            double r_c = sqrt(x_c * x_c + y_c * y_c), p_c, zh0, zh02, dum, xh0;
            if (x_c == 0.0) x_c = x_c + 0.001 * spx;
            p_c = atan(y_c / x_c) - phioff;
            if (x_c < 0.0) p_c = p_c + 3.14159265359;
            while (p_c < 0.0) p_c = p_c + 6.28318530718;
            while (p_c >= 6.28318530718) p_c = p_c - 6.28318530718;
            zh0 = -sin(p_c) / sinth0;
            zh02 = zh0 * zh0;
            dum = 1.0 - zh02 * sinth0 * sinth0;
            dum = dum + 1e-4;
            if (dum < 0.0) return 13;
            xh0 = cos(p_c) > 0.0 ? sqrt(dum) : -sqrt(dum);
            rays->x0[ir] = r_c * xh0;
            rays->z0[ir] = r_c * zh0;
            b[ir] = r_c;
            ir++;
        }
    }
    *starunres = 0;
    if (addstar > 0) {                                             /* :2349-2400 */
        if (xoff != 0.0 || yoff != 0.0) return 13;
        rays->x0[ir] = 0.0;
        rays->z0[ir] = 0.0;
        b[ir] = 0.0;
        if (spx * spx + spy * spy > rstar * rstar) *starunres = 1;
    }
    *spx_out = spx;
    *spy_out = spy;
    return 0;
}

/* I (erg/s/cm2/Hz/sr) and the line-of-sight optical depth per pixel and
 * channel, image[iy][ix][ch] and tau[iy][ix][ch].  Every channel is
 * integrated.  The pos-vel path of the reference does not set the star's
 * passband spectrum (calc_write_line_posvel calls line_setup_passband but
 * not line_init_passband_starbc, which calc_line_spectrum calls at
 * telescope.F:1557), so the stellar intensity there is 0. */
int orc_posvel(const orc_model *m, const orc_grid *g, const orc_rays *rays, const double *b,
               int nx, int ny, double spx, double spy, int starunres, int il,
               double width, int nfr, int nthreads, double *image, double *tau,
               int *maser_out, long *overshoot_out)
{
    double nu0 = fabs(m->nu0[il]), passb, nu1, dnu, *dnu_ch, *velo, Bud, Bdu, I_bc = 0.0;
    double rmax = RC(g, g->nr), *cen = NULL;
    int k, err = 0, maser = 0, npix = nx * ny;
    long ovs = 0;
    if (nfr < 2) return 13;
    passb = 3.33567e-6 * nu0 * width;                  /* line.F:462-469 */
    nu1 = 0.0 - passb;
    dnu = 2.0 * passb / (nfr - 1.0);
    dnu_ch = (double *)malloc(sizeof(double) * nfr);
    velo = (double *)malloc(sizeof(double) * nfr);
    for (k = 0; k < nfr; k++) {
        dnu_ch[k] = nu1 + k * dnu;
        velo[k] = dnu_ch[k] / m->nu0[il];
    }
    Bud = 6.78171833781e46 * m->A[il] / (m->nu0[il] * m->nu0[il] * m->nu0[il]);
    Bdu = Bud * (m->gu[il] / m->gl[il]);
    if (m->outer_bc == 2) {
        double f = m->nu0[il];
        I_bc = 1.47455253991e-47 * (f * f * f) / (exp(4.7991598e-11 * f / 2.728) - 1.0);
    }
    if (starunres) cen = (double *)calloc(nfr, sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+:ovs) reduction(|:maser)
#endif
    for (int iray = 0; iray < npix + (starunres ? 1 : 0); iray++) {
        orc_traj t;
        raycache_t cache;
        line_ctx L;
        int e, ch, mz = 0, ix = iray / ny, iy = iray % ny;
        double *row = iray < npix ? image + ((size_t)iy * nx + ix) * nfr : cen;
        double *trow = iray < npix ? tau + ((size_t)iy * nx + ix) * nfr : NULL;
        if (iray < npix && !(b[iray] < 0.999 * rmax)) {   /* outside: black, tau 0 */
            for (ch = 0; ch < nfr; ch++) { row[ch] = 0.0; trow[ch] = 0.0; }
            continue;
        }
        memset(&t, 0, sizeof t);
        memset(&cache, 0, sizeof cache);
        e = orc_trajectory(&t, g, m->rstar, rays->x0[iray], rays->z0[iray], rays->theta0, 0.0);
        if (!e) e = build_cache(m, g, &t, il, 0.0, &cache);
        if (e) {
#ifdef _OPENMP
#pragma omp critical
#endif
            err = e;
            orc_traj_free(&t);
            free_cache(&cache);
            continue;
        }
        L.nu0 = m->nu0[il]; L.A = m->A[il]; L.Bud = Bud; L.Bdu = Bdu;
        for (ch = 0; ch < nfr; ch++) {
            L.dnu_ch = dnu_ch[ch];
            L.u_ch = velo[ch];
            row[ch] = charint(&cache, &L, 0.0, I_bc, &ovs, &mz);
            if (trow) trow[ch] = L.tau;
        }
        if (mz) maser = 1;
        orc_traj_free(&t);
        free_cache(&cache);
    }
    if (!err && starunres) {                                       /* :2153-2200 */
        double srat = (3.14159265 * m->rstar * m->rstar) / (4.0 * spx * spy);
        double srat1 = 1.0 - srat;
        int ixc = nx / 2 - 1, iyc = ny / 2 - 1, dx, dy;
        for (k = 0; k < nfr; k++) {
            double dummy = cen[k] * srat;
            for (dy = 0; dy < 2; dy++)
                for (dx = 0; dx < 2; dx++) {
                    double *p = image + ((size_t)(iyc + dy) * nx + ixc + dx) * nfr + k;
                    *p = dummy + srat1 * *p;
                }
        }
    }
    if (maser_out) *maser_out = maser;
    if (overshoot_out) *overshoot_out = ovs;
    free(dnu_ch); free(velo); free(cen);
    return err;
}
