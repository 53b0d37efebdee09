// Kernels of the raytracer and their launchers.
//
//   k_traj        one thread per ray: trajectory (FP64), node geometry
//                 (segment length, projected velocity, interpolation
//                 stencil), CSR by ray.  Run twice: count, then write.
//   k_cells       per line: cell table {jd, ad, lw, N_up, N_low}
//   k_nodes<T>    one wave per ray: interpolated node records for one
//                 line + the NONREDUNDANT window and continuum channel
//   k_formal<T>   one lane per (ray, channel): formal solution along the
//                 ray (charintline, telescope.F:3889-4312)
//   k_fixup<T>    one thread per ray: the reference's lazy continuum
//                 copies for skipped channels (telescope.F:583-611)
//   k_flux        ring-weighted sums of the image into the spectrum
//                 (calc_freq_flux_observer, telescope.F:1386-1435)
#include "rl_engine.h"

#include <hip/hip_runtime.h>

#include <type_traits>

#include "rl_traj.h"

namespace rl {

// ---------------------------------------------------------------------
// trajectories
// ---------------------------------------------------------------------
struct CellView {
    int nh;                   // nth_half
    const double *vel;        // [cell][3]
    __device__ int cell(int ir, int it) const { return (ir - 1) * nh + (it - 1); }
};

// Stage 1, one thread per ray: the crossing list (s, surface, index) in
// the reference's order, with the hole-chord and central-beam flags, into
// a fixed-size slot per ray (nmax bounds the node count).
// code = icross | rext << 2 | NF_* << 3 | idx << 5
__global__ void __launch_bounds__(128)
k_cross(GridView g, const double *__restrict__ tan2, int nrays,
        const double *__restrict__ x0, const double *__restrict__ z0, double theta0,
        double costh0, double sinth0, int central_blend, double *scr_s1, double *scr_s2,
        int *scr_iy, int nmax, int *__restrict__ count, double *__restrict__ cs,
        int *__restrict__ ccode, int *err) {
    int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    Traj t(g, tan2, x0[ray], z0[ray], theta0, costh0, sinth0, scr_s1 + ray, scr_s2 + ray,
           scr_iy + ray, nrays);
    const long long base = (long long)ray * nmax;
    int icr_prev = -1, idx_prev = -1, star_done = 0;
    const bool central = ray == 0 && central_blend;
    auto sink = [&](int k, const Crossing &c) {
        if (k >= nmax) { set_err(err, E_CAP); return; }
        int flag = 0;
        if (k > 0) {
            // hole chord (telescope.F:4095-4104): two consecutive crossings of R_1
            if (c.icross == 1 && c.idx == 1 && icr_prev == 1 && idx_prev == 1) flag |= NF_RESET;
            // central-beam star blend at the first node on R_1 (telescope.F:4144)
            if (central && star_done == 0 && c.icross == 1 && c.idx == 1) {
                flag |= NF_BLEND;
                star_done = 1;
            }
        }
        cs[base + k] = c.s;
        ccode[base + k] = c.icross | (c.rext << 2) | (flag << 3) | (c.idx << 5);
        icr_prev = c.icross;
        idx_prev = c.idx;
    };
    int n = trajectory(t, sink);
    if (n < 0) { set_err(err, -n); n = 0; }
    if (n > nmax) { set_err(err, E_CAP); n = nmax; }
    count[ray] = n;
}

__device__ __forceinline__ void geo_node(GridView g, CellView cv, int k, double x0, double z0, double costh0,
                         double sinth0, const double *__restrict__ cs,
                         const int *__restrict__ ccode, GeoNode *__restrict__ nodes, int *err);

// Stage 2, one thread per node: position, cell, interpolation stencil,
// local direction (telescope.F:3682-3718) and projected velocity
// (get_line_dust_values + omega_dot_v, line.F:4039-4208, 2650-2662),
// compacted from the per-ray slots into the CSR node array.
__device__ __forceinline__ void geo_node(GridView g, CellView cv, int k, double x0, double z0, double costh0,
                         double sinth0, const double *__restrict__ cs,
                         const int *__restrict__ ccode, GeoNode *__restrict__ nodes, int *err) {
    const double pi = PICONST, epsplus = 1.0e1 * TELESC_EPS;
    const int nr = g.nr, ny = g.ny;
    const double sinth02 = sinth0 * sinth0, costh02 = costh0 * costh0;
    const double s = cs[k];
    const int code = ccode[k];
    const int icross = code & 3, rext = (code >> 2) & 1, flag = (code >> 3) & 3, idx = code >> 5;
    auto thr = [&](double q) { return sqrt(x0 * x0 + z0 * z0 + q * q + 2.0 * z0 * costh0 * q); };
    auto theta_of_s = [&](double q) {
        double th = atan(sqrt(x0 * x0 + sinth02 * q * q) / (z0 + costh0 * q));
        if (th < 0.0) th = th + pi;
        return th;
    };
    auto hunt_r = [&](double x) { return hunt_bisect([&](int i) { return g.R(i); }, nr, x); };
    auto hunt_t = [&](double x) { return hunt_bisect([&](int i) { return g.TH(i); }, ny, x); };
    double r, theta;
    int ir, it;
    if (icross == 1) {
        ir = idx;
        r = idx == 0 ? g.rstar : g.R(idx);
        theta = theta_of_s(s);
        it = hunt_t(theta);
    } else if (icross == 2) {
        r = thr(s);
        ir = r < g.R(1) ? 0 : (r > g.R(nr) ? nr : hunt_r(r));
        it = idx;
        theta = g.TH(idx);
    } else {
        if (rext) {
            r = sqrt(x0 * x0 + z0 * z0 * (1.0 - costh02));
            ir = idx;
        } else {
            r = thr(s);
            ir = hunt_r(r);
        }
        theta = theta_of_s(s);
        it = hunt_t(theta);
    }
    GeoNode o;
    if (ir == 0) { set_err(err, E_HOLE); ir = 1; }
    double dr = (r - g.R(ir)) / (g.R(ir + 1) - g.R(ir));
    double dt = (theta - g.TH(it)) / (g.TH(it + 1) - g.TH(it));
    if (dr < 0.0 || dr > 1.0) set_err(err, E_DR);
    if (dt < 0.0 || dt > 1.0) set_err(err, E_DT);
    int r1, t1;
    if (dr > 0.0) { r1 = ir + 1; if (r1 > nr) r1 = nr; }
    else { r1 = ir - 1; if (r1 < 1) r1 = 1; }
    t1 = dt > 0.0 ? it + 1 : it - 1;
    const int t0 = ridx_it(it, ny);
    t1 = ridx_it(t1, ny);
    o.c00 = cv.cell(ir, t0);
    o.c01 = cv.cell(ir, t1);
    o.c10 = cv.cell(r1, t0);
    o.c11 = cv.cell(r1, t1);
    o.dr = dr;
    o.dt = dt;
    // local direction, telescope.F:3682-3718
    const double znew = z0 * sinth02, bnew = sqrt(x0 * x0 + z0 * z0 * sinth02);
    const double snew = s + z0 * costh0;
    const double mu = snew / sqrt(bnew * bnew + snew * snew);
    const double q = znew + snew * costh0;
    const double dummy = bnew * sqrt(bnew * bnew + snew * snew - q * q);
    double sinphi = dummy > 0.0 ? (bnew * bnew * costh0 - znew * snew) / dummy : epsplus;
    sinphi = x0 < 0.0 ? asin(sinphi) : pi - asin(sinphi);
    while (sinphi < 0.0) sinphi = sinphi + 2.0 * pi;
    while (sinphi >= 2.0 * pi) sinphi = sinphi - 2.0 * pi;
    double v[3];
    const double *V = cv.vel;
    for (int c = 0; c < 3; c++) {
        double a00 = V[o.c00 * 3 + c], a01 = V[o.c01 * 3 + c];
        double a10 = V[o.c10 * 3 + c], a11 = V[o.c11 * 3 + c];
        if (icross == 1) v[c] = (1.0 - dt) * a00 + dt * a01;
        else if (icross == 2) v[c] = (1.0 - dr) * a00 + dr * a10;
        else v[c] = (1.0 - dr) * ((1.0 - dt) * a00 + dt * a01) +
                    dr * ((1.0 - dt) * a10 + dt * a11);
    }
    if (mu > 1.0) set_err(err, 393);
    o.u = 3.335668e-11 * (mu * v[0] + sqrt(1.0 - mu * mu) *
                                         (v[1] * sin(sinphi) + v[2] * cos(sinphi)));
    if (k == 0) {
        o.ds = 0.0;
    } else {
        o.ds = s - cs[k - 1];
        if (o.ds < 0.0) set_err(err, E_DSNEG);
    }
    o.flag = icross | (flag << 2);
    nodes[k] = o;
}

__global__ void __launch_bounds__(256)
k_geo(GridView g, CellView cv, int nmax, const long long *__restrict__ off,
      const double *__restrict__ x0a, const double *__restrict__ z0a, double costh0,
      double sinth0, const double *__restrict__ cs_all, const int *__restrict__ ccode_all,
      GeoNode *__restrict__ nodes_all, int *err) {
    const int ray = blockIdx.x;
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    const long long o = off[ray];
    if (k >= (int)(off[ray + 1] - o)) return;
    geo_node(g, cv, k, x0a[ray], z0a[ray], costh0, sinth0, cs_all + (long long)ray * nmax,
             ccode_all + (long long)ray * nmax, nodes_all + o, err);
}

// ---------------------------------------------------------------------
// per-line cell tables: N = pop*abund*rho*molpg (line.F:4039-4048)
// ---------------------------------------------------------------------
__global__ void k_cells(int ncell, int nlev, int nlines, int line, int lev_up,
                        int lev_low, double molpg, const double *__restrict__ rho,
                        const double *__restrict__ abund, const double *__restrict__ lw,
                        const double *__restrict__ pops, const double *__restrict__ jd,
                        const double *__restrict__ ad, CellRec *__restrict__ out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    CellRec r;
    r.v[CR_JD] = jd[(size_t)c * nlines + line];
    r.v[CR_AD] = ad[(size_t)c * nlines + line];
    r.v[CR_LW] = lw[c];
    r.v[CR_NU] = pops[(size_t)c * nlev + lev_up] * abund[c] * rho[c] * molpg;
    r.v[CR_NL] = pops[(size_t)c * nlev + lev_low] * abund[c] * rho[c] * molpg;
    out[c] = r;
}

// ---------------------------------------------------------------------
// node records (one wave per ray) + window + continuum channel
// ---------------------------------------------------------------------
// get_line_dust_values interpolation (line.F:4006-4208) of the cell
// table {jd, ad, lw, N_up, N_low} at one node
__device__ inline void node_interp(const GeoNode &gn, const CellRec *__restrict__ cells,
                                   double v[5]) {
    const CellRec c00 = cells[gn.c00], c01 = cells[gn.c01];
    const CellRec c10 = cells[gn.c10], c11 = cells[gn.c11];
    const double dr = gn.dr, dt = gn.dt;
    const int icr = gn.flag & 3;
    for (int f = 0; f < 5; f++) {
        if (icr == 1) v[f] = (1.0 - dt) * c00.v[f] + dt * c01.v[f];
        else if (icr == 2) v[f] = (1.0 - dr) * c00.v[f] + dr * c10.v[f];
        else v[f] = (1.0 - dr) * ((1.0 - dt) * c00.v[f] + dt * c01.v[f]) +
                    dr * ((1.0 - dt) * c10.v[f] + dt * c11.v[f]);
    }
}

// NONREDUNDANT window and first skipped channel of one ray (one wave)
__device__ inline void finish_window(int ray, int lane, double umin, double umax, int nfr,
                                     const double *__restrict__ velo, double aks2,
                                     int skip_mode, int n, RayInfo *__restrict__ rinfo) {
    for (int m = 32; m >= 1; m >>= 1) {
        umin = fmin(umin, __shfl_xor(umin, m));
        umax = fmax(umax, __shfl_xor(umax, m));
    }
    float mn = umin < 1.0 ? (float)umin : 1.0f;
    float mx = umax > -1.0 ? (float)umax : -1.0f;
    int cont = nfr;
    if (skip_mode == 0 && ray > 0) {
        for (int ch = lane; ch < nfr; ch += 64) {
            bool inwin = velo[ch] <= (double)mx + aks2 && velo[ch] >= (double)mn - aks2;
            if (!inwin) { cont = ch; break; }
        }
        for (int m = 32; m >= 1; m >>= 1) cont = min(cont, __shfl_xor(cont, m));
    }
    if (lane == 0) {
        RayInfo ri;
        ri.mn = mn; ri.mx = mx; ri.cont = cont < nfr ? cont : -1; ri.n = n;
        rinfo[ray] = ri;
    }
}

// FP64 parity path: node values as the reference holds them
__global__ void __launch_bounds__(64)
k_nodes(int nrays, const long long *__restrict__ off, const GeoNode *__restrict__ geo,
        const CellRec *__restrict__ cells, NodeRec<double> *__restrict__ rec,
        RayInfo *__restrict__ rinfo, int nfr, const double *__restrict__ velo,
        double aks2, int skip_mode) {
    const int ray = blockIdx.x;
    const int lane = threadIdx.x;
    if (ray >= nrays) return;
    const long long b = off[ray];
    const int n = (int)(off[ray + 1] - b);
    double umin = 1.0e300, umax = -1.0e300;
    for (int k = lane; k < n; k += 64) {
        const GeoNode gn = geo[b + k];
        double v[5];
        node_interp(gn, cells, v);
        NodeRec<double> r;
        r.ds = gn.ds; r.u = gn.u;
        r.jd = v[CR_JD]; r.ad = v[CR_AD]; r.lw = v[CR_LW]; r.nu = v[CR_NU]; r.nl = v[CR_NL];
        r.flag = (double)(gn.flag >> 2);
        rec[b + k] = r;
        // window from the node preceding each segment (telescope.F:4265-4270)
        if (k + 1 < n && v[CR_NU] + v[CR_NL] > LEVTHRES) {
            umin = fmin(umin, gn.u);
            umax = fmax(umax, gn.u);
        }
    }
    finish_window(ray, lane, umin, umax, nfr, velo, aks2, skip_mode, n, rinfo);
}

// per-node scale for the FP32 records: the dust opacity where there is
// dust, else the largest line coefficient, else 1
__device__ inline void scaled_node(const double v[5], const LineConst &lc, double &sig,
                                   double &jc, double &dc) {
    jc = lc.hknu * v[CR_NU] * lc.A;
    dc = lc.hknu * (v[CR_NL] * lc.Bdu - v[CR_NU] * lc.Bud);
    sig = v[CR_AD];
    if (!(sig > 0.0)) sig = fmax(fmax(fabs(dc), fabs(jc)), fabs(v[CR_JD]));
    if (!(sig > 0.0)) sig = 1.0;
}

// FP32 performance path: channel-independent products precomputed
__global__ void __launch_bounds__(64)
k_nodes_fast(int nrays, const long long *__restrict__ off, const GeoNode *__restrict__ geo,
             const CellRec *__restrict__ cells, FastRec *__restrict__ rec,
             RayInfo *__restrict__ rinfo, int nfr, const double *__restrict__ velo,
             double aks2, int skip_mode, LineConst lc) {
    const int ray = blockIdx.x;
    const int lane = threadIdx.x;
    if (ray >= nrays) return;
    const long long b = off[ray];
    const int n = (int)(off[ray + 1] - b);
    double umin = 1.0e300, umax = -1.0e300;
    for (int k = lane; k < n; k += 64) {
        const GeoNode gn = geo[b + k];
        double v[5];
        node_interp(gn, cells, v);
        FastRec r;
        double sig, jc, dc;
        scaled_node(v, lc, sig, jc, dc);
        r.ds = (float)gn.ds;
        r.u = (float)gn.u;
        r.sig = (float)sig;
        r.jdh = (float)(v[CR_JD] / sig);
        r.adh = (float)(v[CR_AD] / sig);
        r.jch = (float)(jc / sig);
        r.dch = (float)(dc / sig);
        r.flag = (float)(gn.flag >> 2);
        r.pad = 0.f;
        if (k > 0) {
            const GeoNode gp = geo[b + k - 1];
            double vp[5], sp, jcp, dcp;
            node_interp(gp, cells, vp);
            scaled_node(vp, lc, sp, jcp, dcp);
            double lwav = 0.5 * (vp[CR_LW] + v[CR_LW]);
            double aa = lc.c_aa * lwav;
            r.inv_aa = (float)(1.0 / aa);
            r.norm = (float)(0.56419583546 / aa);
            double du = gn.u - gp.u;
            double x = fabs(du / (lwav / 2.99792458e5));
            r.dls = 2.0 * 3.0 * x > 1.0 ? (float)(gn.ds / x) : -1.0f;
            r.inv_du = (float)(1.0 / du);
            double sm = fmax(sp, sig);
            r.rel0 = (float)(sp / sm);
            r.rel1 = (float)(sig / sm);
            r.sigm = (float)sm;
        } else {
            r.inv_aa = 0.f; r.norm = 0.f; r.dls = -1.f; r.inv_du = 0.f;
            r.rel0 = r.rel1 = 1.f; r.sigm = r.sig;
        }
        rec[b + k] = r;
        if (k + 1 < n && v[CR_NU] + v[CR_NL] > LEVTHRES) {
            umin = fmin(umin, gn.u);
            umax = fmax(umax, gn.u);
        }
    }
    finish_window(ray, lane, umin, umax, nfr, velo, aks2, skip_mode, n, rinfo);
}

// ---------------------------------------------------------------------
// formal solution
// ---------------------------------------------------------------------
template <typename T>
__device__ inline T trace_one(const NodeRec<T> *__restrict__ rec, int n, const LineK<T> &L,
                              T I0, T istar, T starfract, int &ovs, int &maser) {
    T I = I0;
    Carry<T> c;
    c.init = 1; c.phi0 = T(0); c.srcl0 = T(0); c.alpl0 = T(0); c.S0 = T(0);
    NodeRec<T> a = rec[0];
    Pt<T> pa{a.jd, a.ad, a.u, a.nu, a.nl};
    T lwa = a.lw;
    for (int k = 1; k < n; k++) {
        NodeRec<T> b = rec[k];
        T ds = b.ds;
        int fl = (int)b.flag;
        if (fl & NF_RESET) { ds = T(0); c.init = 1; }
        if (fl & NF_BLEND) {
            I = (T(1) - starfract) * I + starfract * istar;
            c.init = 1;
        }
        Pt<T> pb{b.jd, b.ad, b.u, b.nu, b.nl};
        clever_elem<T>(L, I, ds, pa, pb, lwa, b.lw, c, ovs);
        if (c.alpl0 * ds < T(-0.01)) maser = 1;
        pa = pb;
        lwa = b.lw;
    }
    return I;
}

template <typename T> struct RecOf { using type = NodeRec<double>; };
template <> struct RecOf<float> { using type = FastRec; };

template <typename T>
__device__ inline double trace_dispatch(const typename RecOf<T>::type *rec, long long b, int n,
                                        const LineConst &lc, double dnu, double velo,
                                        double istar, T sf, int &ovs, int &maser);

template <>
__device__ inline double trace_dispatch<double>(const NodeRec<double> *rec, long long b, int n,
                                                const LineConst &lc, double dnu, double velo,
                                                double istar, double sf, int &ovs,
                                                int &maser) {
    LineK<double> L;
    L.nu0 = lc.nu0; L.c_aa = lc.c_aa; L.hknu = lc.hknu;
    L.A = lc.A; L.Bud = lc.Bud; L.Bdu = lc.Bdu;
    L.dnu = dnu; L.uch = velo;
    return trace_one<double>(rec + b, n, L, lc.I0, istar, sf, ovs, maser);
}

template <>
__device__ inline double trace_dispatch<float>(const FastRec *rec, long long b, int n,
                                               const LineConst &lc, double dnu, double velo,
                                               double istar, float sf, int &ovs, int &maser) {
    FastLine L{(float)lc.nu0, (float)dnu, (float)velo};
    return (double)trace_fast(rec + b, n, L, (float)lc.I0, (float)istar, sf, ovs, maser);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_formal(int nrays, int nfr, const long long *__restrict__ off,
         const typename RecOf<T>::type *__restrict__ rec, const RayInfo *__restrict__ rinfo,
         const double *__restrict__ dnu, const double *__restrict__ velo,
         const double *__restrict__ istar, LineConst lc, double aks2, int skip_mode,
         double *__restrict__ image, unsigned long long *ovs_out, int *maser_out) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int ovs = 0, maser = 0;
    if (g < (long long)nrays * nfr) {
        const int ray = (int)(g / nfr), ch = (int)(g - (long long)ray * nfr);
        const RayInfo ri = rinfo[ray];
        bool doit = true;
        if (skip_mode == 0 && ray > 0 && ch > 0 && ch != ri.cont) {
            doit = velo[ch] <= (double)ri.mx + aks2 && velo[ch] >= (double)ri.mn - aks2;
        }
        double out;
        if (doit) {
            T sf = ray == 0 ? (T)lc.starfract : T(0);
            out = trace_dispatch<T>(rec, off[ray], ri.n, lc, dnu[ch], velo[ch], istar[ch], sf,
                                    ovs, maser);
        } else {
            out = __longlong_as_double(0x7ff8dead00000000ll);   // filled by k_fixup
        }
        image[g] = out;
    }
    // wave-level pre-reduction of the warning counters
    for (int m = 32; m >= 1; m >>= 1) {
        ovs += __shfl_xor(ovs, m);
        maser |= __shfl_xor(maser, m);
    }
    if ((threadIdx.x & 63) == 0) {
        if (ovs) atomicAdd(ovs_out, (unsigned long long)ovs);
        if (maser) atomicOr(maser_out, 1);
    }
}

// lazy continuum (telescope.F:583-611): the first skipped channel of a
// ray was integrated by k_formal; later skipped channels copy it, unless
// it is exactly 0, in which case they are integrated in order.
template <typename T>
__global__ void k_fixup(int nrays, int nfr, const long long *__restrict__ off,
                        const typename RecOf<T>::type *__restrict__ rec, const RayInfo *__restrict__ rinfo,
                        const double *__restrict__ dnu, const double *__restrict__ velo,
                        const double *__restrict__ istar, LineConst lc, double aks2,
                        double *__restrict__ image, unsigned long long *ovs_out,
                        int *maser_out) {
    int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray == 0 || ray >= nrays) return;
    const RayInfo ri = rinfo[ray];
    if (ri.cont < 0) return;
    double *row = image + (size_t)ray * nfr;
    double cont = row[ri.cont];
    int ovs = 0, maser = 0;
    for (int ch = ri.cont + 1; ch < nfr; ch++) {
        bool inwin = velo[ch] <= (double)ri.mx + aks2 && velo[ch] >= (double)ri.mn - aks2;
        if (inwin) continue;
        if (cont != 0.0) {
            row[ch] = cont;
        } else {
            row[ch] = trace_dispatch<T>(rec, off[ray], ri.n, lc, dnu[ch], velo[ch], istar[ch],
                                        T(0), ovs, maser);
            cont = row[ch];
        }
    }
    if (ovs) atomicAdd(ovs_out, (unsigned long long)ovs);
    if (maser) atomicOr(maser_out, 1);
}

// ---------------------------------------------------------------------
// flux: partial ring sums, then a fixed-order sum of the partials
// ---------------------------------------------------------------------
__global__ void k_flux_partial(int nfr, int nrings, int nphi, int rings_per_chunk,
                               const double *__restrict__ image,
                               const double *__restrict__ ring_w, double center_w,
                               double *__restrict__ partial) {
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (ch >= nfr) return;
    double acc = 0.0;
    if (chunk == 0) acc = center_w * image[ch];
    int r0 = chunk * rings_per_chunk, r1 = min(nrings, r0 + rings_per_chunk);
    for (int ir = r0; ir < r1; ir++) {
        const double *row = image + (size_t)(1 + ir * nphi) * nfr + ch;
        double s = 0.0;
        for (int ip = 0; ip < nphi; ip++) s = s + row[(size_t)ip * nfr];
        s = s / (1.0 * nphi);
        acc = acc + s * ring_w[ir];
    }
    partial[(size_t)chunk * nfr + ch] = acc;
}

__global__ void k_flux_final(int nfr, int nchunks, const double *__restrict__ partial,
                             double d2, double *__restrict__ flux) {
    int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= nfr) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s = s + partial[(size_t)c * nfr + ch];
    flux[ch] = s / d2;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
void launch_cross(hipStream_t st, const GridDev &gd, const double *tan2, int nrays,
                  const double *x0, const double *z0, double theta0, double costh0,
                  double sinth0, int central_blend, double *s1, double *s2, int *iy, int nmax,
                  int *count, double *cs, int *ccode, int *err) {
    GridView g{gd.nr, gd.ny, gd.rc, gd.tc, gd.rstar};
    dim3 blk(128), grd((nrays + 127) / 128);
    hipLaunchKernelGGL(k_cross, grd, blk, 0, st, g, tan2, nrays, x0, z0, theta0, costh0,
                       sinth0, central_blend, s1, s2, iy, nmax, count, cs, ccode, err);
}

void launch_geo(hipStream_t st, const GridDev &gd, const double *vel, int nh, int nrays,
                int nmax, const long long *off, const double *x0, const double *z0,
                double costh0, double sinth0, const double *cs, const int *ccode,
                GeoNode *nodes, int *err) {
    GridView g{gd.nr, gd.ny, gd.rc, gd.tc, gd.rstar};
    CellView cv{nh, vel};
    hipLaunchKernelGGL(k_geo, dim3(nrays, (nmax + 255) / 256), dim3(256), 0, st, g, cv, nmax, off,
                       x0, z0, costh0,
                       sinth0, cs, ccode, nodes, err);
}

void launch_cells(hipStream_t st, int ncell, int nlev, int nlines, int line, int lev_up,
                  int lev_low, double molpg, const double *rho, const double *abund,
                  const double *lw, const double *pops, const double *jd, const double *ad,
                  CellRec *out) {
    hipLaunchKernelGGL(k_cells, dim3((ncell + 255) / 256), dim3(256), 0, st, ncell, nlev,
                       nlines, line, lev_up, lev_low, molpg, rho, abund, lw, pops, jd, ad,
                       out);
}

template <typename T>
void launch_item(hipStream_t st, const ItemLaunch &it) {
    auto *rec = reinterpret_cast<typename RecOf<T>::type *>(it.rec);
    if constexpr (std::is_same<T, double>::value)
        hipLaunchKernelGGL(k_nodes, dim3(it.nrays), dim3(64), 0, st, it.nrays, it.off,
                           it.geo, it.cells, rec, it.rinfo, it.nfr, it.velo, it.aks2,
                           it.skip_mode);
    else
        hipLaunchKernelGGL(k_nodes_fast, dim3(it.nrays), dim3(64), 0, st, it.nrays, it.off,
                           it.geo, it.cells, rec, it.rinfo, it.nfr, it.velo, it.aks2,
                           it.skip_mode, it.lc);
    if (it.ev_nodes) (void)hipEventRecord(it.ev_nodes, st);
    long long total = (long long)it.nrays * it.nfr;
    hipLaunchKernelGGL(k_formal<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       it.nrays, it.nfr, it.off, rec, it.rinfo, it.dnu, it.velo, it.istar,
                       it.lc, it.aks2, it.skip_mode, it.image, it.ovs, it.maser);
    if (it.skip_mode == 0)
        hipLaunchKernelGGL(k_fixup<T>, dim3((it.nrays + 127) / 128), dim3(128), 0, st,
                           it.nrays, it.nfr, it.off, rec, it.rinfo, it.dnu, it.velo,
                           it.istar, it.lc, it.aks2, it.image, it.ovs, it.maser);
    if (it.ev_formal) (void)hipEventRecord(it.ev_formal, st);
    dim3 gp((it.nfr + 63) / 64, it.nchunks);
    hipLaunchKernelGGL(k_flux_partial, gp, dim3(64), 0, st, it.nfr, it.nrings, it.nphi,
                       it.rings_per_chunk, it.image, it.ring_w, it.center_w, it.partial);
    hipLaunchKernelGGL(k_flux_final, dim3((it.nfr + 63) / 64), dim3(64), 0, st, it.nfr,
                       it.nchunks, it.partial, it.d2, it.flux);
}

template void launch_item<double>(hipStream_t, const ItemLaunch &);
template void launch_item<float>(hipStream_t, const ItemLaunch &);

}  // namespace rl
