/*
 * CPU FP64 restatement of RADLite's circular-image line raytrace.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the
 * checker or as the timed CPU baseline -- never as a product path.
 *
 * Parity: pinned against the compiled reference (oracle/refbuild) on the
 * golden fixtures in tests/golden (per-ray, per-channel intensities and
 * spectra) and against closed forms in tests/test_oracle_*.py.
 *
 * Array conventions: cell fields are [nr][nth_half] with theta fastest
 * (the order of the reference readers, setup.F:778-785); populations
 * [nr][nth_half][nlev]; per-line dust tables [nr][nth_half][nlines].
 * All line/level indices passed in are 0-based.
 */
#ifndef RL_ORACLE_H
#define RL_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nr, nth_half;
    const double *r;          /* [nr]      radius.inp                    */
    const double *theta_half; /* [nth_half] theta.inp                    */
    const double *rho, *abund, *vel, *lw;   /* cells; vel has 3 comps   */
    int nlev;
    const double *pops;       /* [nr][nth][nlev]                         */
    double umass_av;
    int nlines;
    const double *nu0, *A, *gu, *gl;        /* per line                  */
    const int *lev_up, *lev_low;            /* 0-based level indices     */
    const double *jd, *ad;    /* [nr][nth][nlines] dust j, alpha at nu0  */
    double aksmax;            /* km/s, line.F:3020-3033                  */
    double rstar;
    int ncont;
    const double *cont_nu, *cont_istar;     /* star I_nu at R*           */
    int outer_bc;             /* 0 = black, 2 = CMB, 3 = interstellar    */
    const double *cont_isrf;  /* [ncont] interstellfield.inp (type 3)    */
} orc_model;

typedef struct {
    int nrays, nrings, nphi;
    double theta0;
    double *x0, *z0;          /* [nrays]                                 */
    double *ring_r;           /* [nrings+1], ring_r[0] = 0               */
    double *ring_ri;          /* [nrings+2] image ring edges             */
} orc_rays;

typedef struct {
    int n;
    double *s, *r, *theta, *mu, *phi;
    int *icross, *ir, *it;
    double b;
} orc_traj;

/* Grid with ghosts: rc[i+1] = R_i for i=-1..nr+2; tc likewise. */
typedef struct {
    int nr, nth;              /* nth = full (mirrored) count             */
    double *rc, *tc;
} orc_grid;

int  orc_grid_build(orc_grid *g, int nr, const double *r, int nth_half,
                    const double *theta_half);
void orc_grid_free(orc_grid *g);

int  orc_setup_rays(orc_rays *rays, const orc_grid *g, double incl_rad,
                    int nphi, int nrext, int dbdr, double rstar);
void orc_rays_free(orc_rays *rays);

int  orc_trajectory(orc_traj *t, const orc_grid *g, double rstar,
                    double x0, double z0, double theta0, double rs0_fill);
void orc_traj_free(orc_traj *t);

/* One line at one inclination (calc_line_spectrum, telescope.F:1465).
 * skip_mode 0 = the reference's NONREDUNDANT channel skipping,
 * 1 = integrate every channel.  image: [nrays][nfr] or NULL.
 * ray_stride > 1 traces only every ray_stride-th ray (image rows of the
 * others are left untouched and the flux is not computed). 
 * Returns 0 or a reference stop code. */
int orc_line_spectrum(const orc_model *m, const orc_grid *g,
                      const orc_rays *rays, int iline, double vmax_kms,
                      int nfr, int skip_mode, int nthreads, int ray_stride,
                      double *flux, double *image, int *cmask,
                      float *minvel, float *maxvel, double *dnu_out,
                      int *maser, long *overshoot, long *n_integrated);

/* Host table builders with the reference's arithmetic (R13). */
/* Rectangular image rays (setup_rays_rectang, telescope.F:2229-2468):
 * pixel rays ix outer, iy inner, then the central ray when addstar;
 * b[] receives each ray's pixel radius (rp_b).  sizes: the whole image. */
int orc_setup_rays_rect(orc_rays *rays, double *b, double anginf, int nx, int ny,
                        double size_x, double size_y, double phioff, double xoff,
                        double yoff, int addstar, double rstar, double *spx_out,
                        double *spy_out, int *starunres);
/* Position-velocity cube of line il (make_freq_image_rectang,
 * telescope.F:2061-2183): image and tau [ny][nx][nfr]. */
int orc_posvel(const orc_model *m, const orc_grid *g, const orc_rays *rays, const double *b,
               int nx, int ny, double spx, double spy, int starunres, int il,
               double width, int nfr, int nthreads, double *image, double *tau,
               int *maser_out, long *overshoot_out);
double orc_bplanck(double temp, double nu);
void orc_line_consts(int nlev, const double *E_cm, int nlines,
                     const int *lev_up, const int *lev_low,
                     double *nu0_out);
/* scat: the scattering/PAH source [nfreq][nr][nth_half] (scatsource.dat
 * + quantsource.dat) or NULL */
int  orc_line_dust(int nr, int nth_half, int nspec, const double *rho_d,
                   const double *t_d, int nfreq, const double *cont_nu,
                   const double *kabs, const double *kscat, const double *scat,
                   int nlines, const double *nu0, double *jd, double *ad);
/* Outer boundary intensity of the channels of a passband (charintline,
 * telescope.F:3996-4026): 0, the CMB at the line centre, or the
 * interstellar field at each channel frequency
 * (line_init_passband_outerbc, line.F:3855-3902). */
void orc_outer_bc(const orc_model *m, int il, int nfr, const double *dnu_ch, double *I_bc);
void orc_star_intensity(int nfreq, const double *fnu_1pc, double rstar,
                        double *istar);
double orc_aksmax(int ncell, const double *lw, double fix_lw);
int  orc_passband_nfr(double vmax, double dv, double *vmax_out);

/* Test hook: node values of one ray (ds[0] carries starfract for ray 0). */
int orc_ray_nodes(const orc_model *m, const orc_grid *g, const orc_rays *rays, int il,
                  int iray, int cap, double *ds, double *u, double *jd, double *ad,
                  double *lw, double *nu, double *nl, int *flags);

/* Exposed pieces for unit tests. */
double orc_qdr_src_2(double I, double js1, double alp1, double js2,
                     double alp2, double ds, long *overshoot);
double orc_voigt(double nu0, double aks, double dnu);
int    orc_hunt(const double *xx, int n, double x, int jlo);

#ifdef __cplusplus
}
#endif
#endif
