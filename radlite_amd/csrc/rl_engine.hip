// Kernels of the raytracer and their launchers.
//
//   k_cross       one thread per ray: the crossings of the trajectory
//                 (make_trajectory_c, telescope.F:2787-3677) in FP64,
//                 written into per-ray slots
//   k_geo         one thread per node: position, stencil, local direction
//                 and projected velocity (telescope.F:3682-3718,
//                 line.F:2650-2662), compacted into the CSR node array
//   k_cells       per line: cell table {jd, ad, lw, N_up, N_low}
//   k_nodes       one wave per ray: interpolated node records (FP64) +
//                 the NONREDUNDANT window and continuum channel
//   k_nodes_fast  the same as scaled records (FP32 mode, adjoint)
//   k_formal_batch<T> / k_formal_tiles<T>  one lane per (ray, channel) of
//                 many items: formal solution along the ray (charintline,
//                 telescope.F:3889-4312)
//   k_fixup_batch<T>  one thread per ray: the reference's lazy continuum
//                 copies for skipped channels (telescope.F:583-611)
//   k_flux_*      ring-weighted sums of the image into the spectrum
//                 (calc_freq_flux_observer, telescope.F:1386-1435)
//   k_adj_*       adjoint: checkpointed forward, reverse sweep, node and
//                 cell gradient scatter
#include "rl_engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "rl_traj.h"
#include "rl_adjoint.h"
#include "rl_fwd32.h"

namespace rl {

// ---------------------------------------------------------------------
// trajectories
// ---------------------------------------------------------------------
struct CellView {
    int nh;                   // nth_half
    const double *vel;        // [cell][3]
    const VelLw *vlw;         // [cell]: velocity and width together (the FP32 node pass)
    __device__ int cell(int ir, int it) const { return (ir - 1) * nh + (it - 1); }
};

// Stage 1, one thread per ray: the crossing list (s, surface, index) in
// the reference's order, with the hole-chord and central-beam flags, into
// a fixed-size slot per ray (nmax bounds the node count).
// code = icross | rext << 2 | NF_* << 3 | idx << 5
// Several ray sets (inclinations) go in one launch, blockIdx.y choosing the
// set: a set has one thread per ray, a long serial chain of FP64 work, so
// sets launched together keep more waves resident to hide its latency.
// A thread's records go through LDS and out 4 at a time, 64 aligned bytes
// of its slot (slots start on 128-byte lines: nmax is a multiple of 8).
// Written one 16-byte record per crossing, the lines of the ~34 MB of slots
// being filled at once were evicted part-written: 4.8 times the records'
// bytes were written, and about as many read back.  Eight records (whole
// lines) would need 16 KB of LDS per block, which with the grid tables
// leaves too few blocks per CU.  Four waves per SIMD (128 registers) keep
// the merge's state out of scratch; the few blocks that do not fit at once
// are the last, the outermost rings' short rays, which start as the first
// blocks finish (five waves, with per-step spill reloads, were slower).
constexpr int CROSS_BLOCK = 128, CROSS_LINE = 4;

__global__ void __launch_bounds__(CROSS_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_cross(GridView g, const double *__restrict__ tan2, CrossSets sets, int nmax, int *err,
        int grid_in_lds) {
    __shared__ CrossRec stage[CROSS_LINE][CROSS_BLOCK];
    // the radii, cone angles and tan^2 table (a few KB) staged in LDS: the
    // merge reads them at every step
    extern __shared__ double lds_grid[];
    if (grid_in_lds) {
        const int nrc = g.nr + 4, ntc = g.ny + 4, nt2 = g.ny / 2 + 1;
        for (int i = threadIdx.x; i < nrc; i += blockDim.x) lds_grid[i] = g.rc[i];
        for (int i = threadIdx.x; i < ntc; i += blockDim.x) lds_grid[nrc + i] = g.tc[i];
        for (int i = threadIdx.x; i < nt2; i += blockDim.x) lds_grid[nrc + ntc + i] = tan2[i];
        __syncthreads();
        g.rc = lds_grid;
        g.tc = lds_grid + nrc;
        tan2 = lds_grid + nrc + ntc;
    }
    const CrossSet &q = sets.s[blockIdx.y];
    const int nrays = q.nrays;
    const double *__restrict__ x0 = q.x0;
    const double *__restrict__ z0 = q.z0;
    const double theta0 = q.theta0, costh0 = q.costh0, sinth0 = q.sinth0;
    const int central_blend = q.central_blend;
    double *scr_s1 = q.s1, *scr_s2 = q.s2;
    int *scr_iy = q.iy;
    int *__restrict__ count = q.count;
    CrossRec *__restrict__ cr = q.cr;
    int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    if (q.bpix && !(q.bpix[ray] < q.bmax)) {     // telescope.F:2118: outside, not traced
        count[ray] = 0;
        return;
    }
    Traj t(g, tan2, x0[ray], z0[ray], theta0, costh0, sinth0, scr_s1 + ray, scr_s2 + ray,
           scr_iy + ray, nrays);
    const long long base = (long long)ray * nmax;
    // the sink's state, packed (the merge holds many values per lane):
    // bit 0 of st, the previous crossing was on R_1; bit 1, the central
    // beam's star blend is still to come; gs, the cell guesses of the next
    // record (the last sphere and cone crossed, plus one, 16 bits each; 0:
    // none or beyond 16 bits)
    int st = ray == 0 && central_blend ? 2 : 0;
    int gs = 0;
    const int tid = threadIdx.x;
    auto flush = [&](int k0, int m) {            // records k0 .. k0+m-1 from the stage
        for (int j = 0; j < m; j++)
            if (k0 + j < nmax) cr[base + k0 + j] = stage[j][tid];
    };
    auto sink = [&](int k, const Crossing &c) {
        if (k >= nmax) { set_err(err, E_CAP); return; }
        const bool on_r1 = c.icross == 1 && c.idx == 1;
        int flag = 0;
        if (k > 0) {
            // hole chord (telescope.F:4095-4104): two consecutive crossings of R_1
            if (on_r1 && (st & 1)) flag |= NF_RESET;
            // central-beam star blend at the first node on R_1 (telescope.F:4144)
            if (on_r1 && (st & 2)) {
                flag |= NF_BLEND;
                st &= ~2;
            }
        }
        stage[k & (CROSS_LINE - 1)][tid] =
            CrossRec{c.s, c.icross | (c.rext << 2) | (flag << 3) | (c.idx << 5), gs};
        if ((k & (CROSS_LINE - 1)) == CROSS_LINE - 1) flush(k - (CROSS_LINE - 1), CROSS_LINE);
        st = (st & ~1) | (on_r1 ? 1 : 0);
        const int v = c.idx + 1 <= 0xFFFF ? c.idx + 1 : 0;
        if (c.icross == 1) gs = (gs & (int)0xFFFF0000) | v;
        else if (c.icross == 2) gs = (gs & 0xFFFF) | (v << 16);
    };
    int n = trajectory(t, sink);
    if (n < 0) { set_err(err, -n); n = 0; }
    if (n > nmax) { set_err(err, E_CAP); n = nmax; }
    flush(n & ~(CROSS_LINE - 1), n & (CROSS_LINE - 1));
    count[ray] = n;
}


// Stage 2, one thread per node: position, cell, interpolation stencil,
// local direction (telescope.F:3682-3718) and projected velocity
// (get_line_dust_values + omega_dot_v, line.F:4039-4208, 2650-2662),
// compacted from the per-ray slots into the CSR node array.
// PACKED: the velocity and the width both from cv.vlw (two 16-byte loads
// per cell instead of four 8-byte ones), the width interpolated into *lwk
// as interp_field forms it; ds is left to the caller, which has the
// previous node's s (*s_out: this node's).
template <bool PACKED>
__device__ __forceinline__ GeoNode geo_node(GridView g, CellView cv, int k, double x0, double z0,
                                            double costh0, double sinth0,
                                            const CrossRec *__restrict__ cr, int *err,
                                            double *lwk = nullptr, double *s_out = nullptr) {
    const double pi = PICONST, epsplus = 1.0e1 * TELESC_EPS;
    const int nr = g.nr, ny = g.ny;
    const double sinth02 = sinth0 * sinth0, costh02 = costh0 * costh0;
    const CrossRec ck = cr[k];
    const double s = ck.s;
    const int code = ck.code;
    const int icross = code & 3, rext = (code >> 2) & 1, flag = (code >> 3) & 3, idx = code >> 5;
    // The position and cells.  A sphere crossing (icross 1) knows its r and
    // radial cell, a cone crossing (2) its theta and theta cell; an extra
    // point (3) knows neither (its r when it is the R-extremum, rext).  The
    // lanes of a wave hold nodes of every kind, so each quantity is formed
    // once for the lanes that need it (one square root for r, one atan for
    // theta, one table search), rather than once per kind in turn.
    const bool need_t = icross != 2, need_r = icross != 1;
    double r = 0.0, theta = 0.0;
    if (icross == 1) r = idx == 0 ? g.rstar : g.R(idx);
    if (icross == 2) theta = g.TH(idx);
    if (need_r) {
        r = sqrt(rext ? x0 * x0 + z0 * z0 * (1.0 - costh02)
                      : x0 * x0 + z0 * z0 + s * s + 2.0 * z0 * costh0 * s);
    }
    if (need_t) {
        theta = atan(sqrt(x0 * x0 + sinth02 * s * s) / (z0 + costh0 * s));
        if (theta < 0.0) theta = theta + pi;
    }
    // the searches (hunt): theta where it is not known, else r (a cone
    // crossing: below R_1 the search gives 0, above R_nr nr); an extra point
    // that is not the R-extremum searches r as well.  The crossing pass's
    // guess, the last sphere / cone crossed before this node, is at most one
    // cell off; a guess that is not is left to the bisection.
    const int gr = (ck.guess & 0xFFFF) - 1, gt = (ck.guess >> 16) - 1;
    int ir = idx, it = idx;
    {
        const bool on_t = need_t;
        const int n = on_t ? ny : nr, gs = on_t ? gt : gr;
        const bool use = gs >= 0 && gs <= n && (!on_t || g.t_sorted);
        const int found = hunt_table(on_t ? g.tc + 1 : g.rc + 1, n, on_t ? theta : r, gs, use);
        if (on_t) it = found;
        else ir = found;
    }
    if (icross == 3 && !rext) ir = hunt_table(g.rc + 1, nr, r, gr, gr >= 0 && gr <= nr);
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
    o.dr = dr;
    o.dt = dt;
    // local direction, telescope.F:3682-3718
    const double znew = z0 * sinth02, bnew = sqrt(x0 * x0 + z0 * z0 * sinth02);
    const double snew = s + z0 * costh0;
    const double mu = snew / sqrt(bnew * bnew + snew * snew);
    const double q = znew + snew * costh0;
    const double dummy = bnew * sqrt(bnew * bnew + snew * snew - q * q);
    // phi = asin(x) (x0 < 0) or pi - asin(x), wrapped to [0, 2 pi); only its
    // sine and cosine are used: sin(phi) = x, cos(phi) = +-sqrt(1 - x^2)
    const double sx = dummy > 0.0 ? (bnew * bnew * costh0 - znew * snew) / dummy : epsplus;
    const double cx = sqrt(1.0 - sx * sx);
    const double sin_phi = sx, cos_phi = x0 < 0.0 ? cx : -cx;
    // the velocity at the node: a sphere or cone crossing interpolates
    // between two cells (in theta, resp. r), an extra point between four
    double v[3];
    const bool two = icross == 1 || icross == 2;
    const int cb = icross == 1 ? o.c01 : o.c10;
    const double w = icross == 1 ? dt : dr;
    if constexpr (PACKED) {
        const VelLw p00 = cv.vlw[o.c00], pb = cv.vlw[cb];
        const double f00[4] = {p00.vx, p00.vy, p00.vz, p00.lw};
        const double fb[4] = {pb.vx, pb.vy, pb.vz, pb.lw};
        double f[4];
        if (two) {
            for (int q = 0; q < 4; q++) f[q] = (1.0 - w) * f00[q] + w * fb[q];
        } else {
            const VelLw p01 = cv.vlw[o.c01], p11 = cv.vlw[o.c11()];
            const double f01[4] = {p01.vx, p01.vy, p01.vz, p01.lw};
            const double f11[4] = {p11.vx, p11.vy, p11.vz, p11.lw};
            for (int q = 0; q < 4; q++)
                f[q] = (1.0 - dr) * ((1.0 - dt) * f00[q] + dt * f01[q]) +
                       dr * ((1.0 - dt) * fb[q] + dt * f11[q]);
        }
        v[0] = f[0]; v[1] = f[1]; v[2] = f[2];
        *lwk = f[3];
    } else {
        const double *V = cv.vel;
        for (int c = 0; c < 3; c++) {
            const double a00 = V[o.c00 * 3 + c], ab = V[cb * 3 + c];
            if (two) {
                v[c] = (1.0 - w) * a00 + w * ab;
            } else {
                const double a01 = V[o.c01 * 3 + c], a11 = V[o.c11() * 3 + c];
                v[c] = (1.0 - dr) * ((1.0 - dt) * a00 + dt * a01) +
                       dr * ((1.0 - dt) * ab + dt * a11);
            }
        }
    }
    if (mu > 1.0) set_err(err, 393);
    o.u = 3.335668e-11 * (mu * v[0] + sqrt(1.0 - mu * mu) *
                                         (v[1] * sin_phi + v[2] * cos_phi));
    if constexpr (PACKED) {
        o.ds = 0.0;
        *s_out = s;
    } else if (k == 0) {
        o.ds = 0.0;
    } else {
        o.ds = s - cr[k - 1].s;
        if (o.ds < 0.0) set_err(err, E_DSNEG);
    }
    o.flag = icross | (flag << 2);
    return o;
}

// ---------------------------------------------------------------------
// split records of the FP32 batched forward (rl_fwd32.h)
// ---------------------------------------------------------------------
// a cell field at a node by the stencil of get_line_dust_values
// (line.F:4006-4037), as node_interp forms it
// (two cells for a sphere or cone crossing, four for an extra point; the
// loads follow, so that a wave of mixed nodes loads two cells per lane)
__device__ inline double interp_field(const GeoNode &gn, const double *__restrict__ f) {
    const double dr = gn.dr, dt = gn.dt;
    const int icr = gn.flag & 3;
    const double a = f[gn.c00], b = f[icr == 1 ? gn.c01 : gn.c10];
    if (icr == 1) return (1.0 - dt) * a + dt * b;
    if (icr == 2) return (1.0 - dr) * a + dr * b;
    return (1.0 - dr) * ((1.0 - dt) * a + dt * f[gn.c01]) + dr * ((1.0 - dt) * b + dt * f[gn.c11()]);
}

// The line-independent records of the segment ending at node gn, from the
// previous node's projected velocity up and width lwp and this node's
// width lwk (first: the ray's first node, which ends no segment).
__device__ inline void seg_record(const GeoNode &gn, bool first, double up, double lwp,
                                  double lwk, SegHot &hot, SegAux &aux) {
    const int fl = gn.flag >> 2;
    SegAux a;
    a.u = (float)gn.u;
    a.flag = fl;
    SegHot h;
    if (first) {
        a.dls = -1.0f; a.inv_du = 0.0f;
        h.hds = -0.0f; h.ag = 0.0f; h.uag = 0.0f; h.ng = 0.0f;
    } else {
        // the segment's mean width (line.F:4704) and its profile scales
        // (voigt_profile line.F:2301-2302 without the line's |nu0|)
        const double lwav = 0.5 * (lwp + lwk);
        const double inv = 1.0 / (3.33567e-6 * lwav);
        const double ag = (double)SQRT_LOG2E * inv;
        const double ds = (fl & NF_RESET) ? 0.0 : gn.ds;     // the hole chord adds nothing
        const double du = gn.u - up;
        const double x = fabs(du / (lwav / 2.99792458e5));   // line.F:4706-4715
        const bool sub = 2.0 * 3.0 * x > 1.0;
        a.dls = sub ? (float)(ds / x) : -1.0f;
        a.inv_du = (float)(1.0 / du);
        const float hds = (float)(0.5 * ds);
        // the sign bit routes the segment to the general path
        h.hds = (sub || fl != 0) ? __int_as_float(__float_as_int(hds) | (int)0x80000000) : hds;
        h.ag = (float)ag;
        h.uag = (float)(gn.u * ag);
        h.ng = (float)(0.56419583546 * inv);
    }
    hot = h;
    aux = a;
}

// GRID_LDS (k_geo, k_geo_seg): the radius and theta tables (with ghosts)
// are staged in LDS, where the cell searches and the cell bounds read them
// without the vector memory path (taken when they fit GEO_GRID_LDS_MAX
// bytes; 13.1 -> 12.45 ms of the 400x200 disk's trajectory build).
constexpr size_t GEO_GRID_LDS_MAX = 12 * 1024;

template <bool GRID_LDS>
__device__ inline void stage_grid(GridView &g, double *lds_tab) {
    const int nrc = g.nr + 4, ntc = g.ny + 4;
    for (int i = threadIdx.x; i < nrc; i += blockDim.x) lds_tab[i] = g.rc[i];
    for (int i = threadIdx.x; i < ntc; i += blockDim.x) lds_tab[nrc + i] = g.tc[i];
    g.rc = lds_tab;
    g.tc = lds_tab + nrc;            // read after the caller's barrier
}

template <bool GRID_LDS>
__global__ void __launch_bounds__(256)
k_geo(GridView g, CellView cv, int nmax, const long long *__restrict__ off,
      const double *__restrict__ x0a, const double *__restrict__ z0a, double costh0,
      double sinth0, const CrossRec *__restrict__ cr_all, GeoNode *__restrict__ nodes_all,
      int *err) {
    extern __shared__ double lds_tab[];
    const int ray = blockIdx.x;
    const long long o = off[ray];
    const int cnt = (int)(off[ray + 1] - o);
    if ((int)(blockIdx.y * blockDim.x) >= cnt) return;       // the whole block: no nodes
    if constexpr (GRID_LDS) {
        stage_grid<GRID_LDS>(g, lds_tab);
        __syncthreads();
    }
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    nodes_all[o + k] = geo_node<false>(g, cv, k, x0a[ray], z0a[ray], costh0, sinth0,
                                cr_all + (long long)ray * nmax, err);
}

// FP32: the nodes and, in the same pass, their segment records, on a grid
// over the ray set's nodes (no block past a ray's last node).  A block
// writes nodes 255 b .. 255 b + 254 of the CSR; its thread 0 forms the node
// before them (not written), whose projected velocity and width the
// block's first segment needs.  A block spans a few rays: its first wave
// finds the first one by a 64-way search of the offsets (four rounds of
// one load per lane for 10^5 rays, where a bisection is eighteen dependent
// loads), the other threads step from it.
template <bool GRID_LDS>
__global__ void __launch_bounds__(256)
k_geo_seg(GridView g, CellView cv, int nmax, int nrays, long long nnodes,
          const long long *__restrict__ off, const RayXZ *__restrict__ xz, double costh0,
          double sinth0, const CrossRec *__restrict__ cr_all, GeoNode *__restrict__ nodes_all,
          SegHot *__restrict__ hot_all, SegAux *__restrict__ aux_all, int *err) {
    __shared__ double s_u[256], s_lw[256], s_s[256];
    __shared__ int s_ray0;
    extern __shared__ double lds_tab[];
    const int t = threadIdx.x;
    if constexpr (GRID_LDS) stage_grid<GRID_LDS>(g, lds_tab);   // read after the barrier below
    const long long base = (long long)blockIdx.x * 255 - 1;       // node of thread 0
    if (t < 64) {
        // the ray of node max(base, 0): the last r with off[r] <= it.  The
        // offsets do not decrease, so the lanes whose candidate passes are
        // a prefix; off[lo] <= g0 and the answer is below hi throughout.
        const long long g0 = base < 0 ? 0 : base;
        int lo = 0, hi = nrays;
        while (hi - lo > 1) {
            const int step = (hi - lo + 63) / 64;
            const int r = lo + t * step;
            const unsigned long long m = __ballot(r < hi && off[r] <= g0);
            lo += (63 - __clzll((long long)m)) * step;
            hi = min(hi, lo + step);
        }
        if (t == 0) s_ray0 = lo;
    }
    __syncthreads();
    const long long gnode = base + t;
    const bool valid = gnode >= 0 && gnode < nnodes;
    GeoNode gn;
    double lwk = 0.0, sk = 0.0;
    int k = 0;
    if (valid) {
        int ray = s_ray0;
        while (off[ray + 1] <= gnode) ray++;
        k = (int)(gnode - off[ray]);
        const RayXZ p = xz[ray];
        gn = geo_node<true>(g, cv, k, p.x0, p.z0, costh0, sinth0,
                            cr_all + (long long)ray * nmax, err, &lwk, &sk);
        s_u[t] = gn.u;
        s_lw[t] = lwk;
        s_s[t] = sk;
    }
    __syncthreads();
    if (!valid || t == 0) return;
    // the path step from the previous node, thread t-1's (the same ray when k > 0)
    if (k > 0) {
        gn.ds = sk - s_s[t - 1];
        if (gn.ds < 0.0) set_err(err, E_DSNEG);
    }
    nodes_all[gnode] = gn;
    seg_record(gn, k == 0, s_u[t - 1], s_lw[t - 1], lwk, hot_all[gnode], aux_all[gnode]);
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
    // a sphere crossing interpolates in theta between c00 and c01, a cone
    // crossing in r between c00 and c10: two cells loaded per lane; only an
    // extra point needs c01 and c11 besides
    const double dr = gn.dr, dt = gn.dt;
    const int icr = gn.flag & 3;
    const CellRec c00 = cells[gn.c00];
    const CellRec cb = cells[icr == 1 ? gn.c01 : gn.c10];
    if (icr == 1 || icr == 2) {
        const double w = icr == 1 ? dt : dr;
        for (int f = 0; f < 5; f++) v[f] = (1.0 - w) * c00.v[f] + w * cb.v[f];
    } else {
        const CellRec c01 = cells[gn.c01], c11 = cells[gn.c11()];
        for (int f = 0; f < 5; f++)
            v[f] = (1.0 - dr) * ((1.0 - dt) * c00.v[f] + dt * c01.v[f]) +
                   dr * ((1.0 - dt) * cb.v[f] + dt * c11.v[f]);
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

// FP64 parity path: node values as the reference holds them (one wave
// per ray)
__device__ inline void nodes_ray_exact(int ray, int lane, const long long *__restrict__ off,
                                       const GeoNode *__restrict__ geo,
                                       const CellRec *__restrict__ cells,
                                       NodeRec<double> *__restrict__ rec,
                                       RayInfo *__restrict__ rinfo, int nfr,
                                       const double *__restrict__ velo, double aks2,
                                       int skip_mode) {
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

__global__ void __launch_bounds__(64)
k_nodes(int nrays, const long long *__restrict__ off, const GeoNode *__restrict__ geo,
        const CellRec *__restrict__ cells, NodeRec<double> *__restrict__ rec,
        RayInfo *__restrict__ rinfo, int nfr, const double *__restrict__ velo,
        double aks2, int skip_mode) {
    if ((int)blockIdx.x >= nrays) return;
    nodes_ray_exact(blockIdx.x, threadIdx.x, off, geo, cells, rec, rinfo, nfr, velo, aks2,
                    skip_mode);
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

// scaled records (FP32 performance path; both precisions of the adjoint):
// channel-independent products precomputed
template <typename T>
__device__ inline void nodes_ray_fast(int ray, int lane, const long long *__restrict__ off,
                                      const GeoNode *__restrict__ geo,
                                      const CellRec *__restrict__ cells,
                                      FastRecT<T> *__restrict__ rec, long long node0,
                                      RayInfo *__restrict__ rinfo, int nfr,
                                      const double *__restrict__ velo, double aks2,
                                      int skip_mode, const LineConst &lc) {
    const long long b = off[ray];
    const int n = (int)(off[ray + 1] - b);
    double umin = 1.0e300, umax = -1.0e300;
    // the previous node's width, scale and projected velocity come from
    // the neighbouring lane (lane 0: the previous chunk's lane 63); every
    // lane runs every chunk so that the shuffles see all of the wave
    double lw_c = 0.0, sig_c = 0.0, u_c = 0.0;
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const bool act = k < n;
        const GeoNode gn = geo[b + (act ? k : n - 1)];
        double v[5];
        node_interp(gn, cells, v);
        double sig, jc, dc;
        scaled_node(v, lc, sig, jc, dc);
        double lw_p = __shfl_up(v[CR_LW], 1), sig_p = __shfl_up(sig, 1), u_p = __shfl_up(gn.u, 1);
        if (lane == 0) { lw_p = lw_c; sig_p = sig_c; u_p = u_c; }
        lw_c = __shfl(v[CR_LW], 63);
        sig_c = __shfl(sig, 63);
        u_c = __shfl(gn.u, 63);
        if (!act) continue;
        FastRecT<T> r;
        r.ds = ((gn.flag >> 2) & NF_RESET) ? T(0) : (T)gn.ds;   // the hole chord adds nothing
        r.u = (T)gn.u;
        r.sig = (T)sig;
        r.jdh = (T)(v[CR_JD] / sig);
        r.adh = (T)(v[CR_AD] / sig);
        r.jch = (T)(jc / sig);
        r.dch = (T)(dc / sig);
        r.flag = (T)(gn.flag >> 2);
        r.pad = T(0);
        if (k > 0) {
            const double sp = sig_p;
            double lwav = 0.5 * (lw_p + v[CR_LW]);
            double aa = lc.c_aa * lwav;
            r.inv_aa = (T)(1.0 / aa);
            r.norm = (T)(0.56419583546 / aa);
            double du = gn.u - u_p;
            double x = fabs(du / (lwav / 2.99792458e5));
            r.dls = 2.0 * 3.0 * x > 1.0 ? (T)(gn.ds / x) : T(-1);
            r.inv_du = (T)(1.0 / du);
            double sm = fmax(sp, sig);
            r.rel0 = (T)(sp / sm);
            r.rel1 = (T)(sig / sm);
            r.sigm = (T)sm;
        } else {
            r.inv_aa = T(0); r.norm = T(0); r.dls = -T(1); r.inv_du = T(0);
            r.rel0 = r.rel1 = T(1); r.sigm = r.sig;
        }
        rec[b - node0 + k] = r;
        if (k + 1 < n && v[CR_NU] + v[CR_NL] > LEVTHRES) {
            umin = fmin(umin, gn.u);
            umax = fmax(umax, gn.u);
        }
    }
    finish_window(ray, lane, umin, umax, nfr, velo, aks2, skip_mode, n, rinfo);
}

// ---------------------------------------------------------------------
// split records of the FP32 batched forward (rl_fwd32.h)
// ---------------------------------------------------------------------
// per node (thread): the line-independent segment records of a ray set
// whose nodes are built (after the widths change)
__global__ void __launch_bounds__(256)
k_segrec(long long nnodes, const long long *__restrict__ off, int nrays,
         const GeoNode *__restrict__ geo, const double *__restrict__ lw,
         SegHot *__restrict__ hot, SegAux *__restrict__ aux) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnodes) return;
    // first node of the ray? (binary search of the CSR)
    int lo = 0, hi = nrays;
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (off[m] <= k) lo = m; else hi = m;
    }
    const bool first = k == off[lo];
    const GeoNode gn = geo[k];
    double up = 0.0, lwp = 0.0;
    if (!first) {
        const GeoNode gp = geo[k - 1];
        up = gp.u;
        lwp = interp_field(gp, lw);
    }
    seg_record(gn, first, up, lwp, interp_field(gn, lw), hot[k], aux[k]);
}

// per item, one wave per ray: the scaled line records and the NONREDUNDANT
// window (telescope.F:4265-4270) of the FP32 batched forward
__device__ inline void nodes_ray_line(int ray, int lane, const long long *__restrict__ off,
                                      const GeoNode *__restrict__ geo,
                                      const CellRec *__restrict__ cells,
                                      NodeLine *__restrict__ rec, RayInfo *__restrict__ rinfo,
                                      int nfr, const double *__restrict__ velo, double aks2,
                                      int skip_mode, const LineConst &lc) {
    const long long b = off[ray];
    const int n = (int)(off[ray + 1] - b);
    const double inv_nu0 = 1.0 / fabs(lc.nu0);
    double umin = 1.0e300, umax = -1.0e300;
    for (int k = lane; k < n; k += 64) {
        const GeoNode gn = geo[b + k];
        double v[5];
        node_interp(gn, cells, v);
        double sig, jc, dc;
        scaled_node(v, lc, sig, jc, dc);
        NodeLine r;
        const float fs = (float)sig;
        r.sig = v[CR_AD] > 0.0 ? fs : __int_as_float(__float_as_int(fs) | (int)0x80000000);
        r.jdh = (float)(v[CR_JD] / sig);
        r.jcl = (float)(jc / sig * inv_nu0);
        r.dcl = (float)(dc / sig * inv_nu0);
        rec[b + k] = r;
        if (k + 1 < n && v[CR_NU] + v[CR_NL] > LEVTHRES) {
            umin = fmin(umin, gn.u);
            umax = fmax(umax, gn.u);
        }
    }
    finish_window(ray, lane, umin, umax, nfr, velo, aks2, skip_mode, n, rinfo);
}

void launch_segrec(hipStream_t st, long long nnodes, const long long *off, int nrays,
                   const GeoNode *geo, const double *lw, SegHot *hot, SegAux *aux) {
    if (nnodes <= 0) return;
    hipLaunchKernelGGL(k_segrec, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, st,
                       nnodes, off, nrays, geo, lw, hot, aux);
}

template <typename T>
__global__ void __launch_bounds__(64)
k_nodes_fast(int nrays, const long long *__restrict__ off, const GeoNode *__restrict__ geo,
             const CellRec *__restrict__ cells, FastRecT<T> *__restrict__ rec, long long node0,
             RayInfo *__restrict__ rinfo, int nfr, const double *__restrict__ velo,
             double aks2, int skip_mode, LineConst lc) {
    if ((int)blockIdx.x >= nrays) return;
    nodes_ray_fast<T>(blockIdx.x, threadIdx.x, off, geo, cells, rec, node0, rinfo, nfr, velo,
                      aks2, skip_mode, lc);
}

// ---------------------------------------------------------------------
// formal solution
// ---------------------------------------------------------------------
template <typename T>
__device__ inline T trace_one(const NodeRec<T> *__restrict__ rec, int n, const LineK<T> &L,
                              T I0, T istar, T starfract, int &ovs, int &maser,
                              T *tau = nullptr) {
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
        if (fl & NF_STAR) {
            I = istar;
            c.init = 1;
        }
        Pt<T> pb{b.jd, b.ad, b.u, b.nu, b.nl};
        clever_elem<T>(L, I, ds, pa, pb, lwa, b.lw, c, ovs);
        if (c.alpl0 * ds < T(-0.01)) maser = 1;
        pa = pb;
        lwa = b.lw;
    }
    if (tau) *tau = c.tau;
    return I;
}

// node records of the batched forward: exact FP64 records, or the FP32
// line records of the split layout (rl_fwd32.h); the position-velocity
// path keeps the whole scaled record per node in FP32
template <typename T> struct RecOf { using type = NodeRec<double>; };
template <> struct RecOf<float> { using type = NodeLine; };
template <typename T> struct PvRecOf { using type = NodeRec<double>; };
template <> struct PvRecOf<float> { using type = FastRec; };

template <typename T>
__device__ inline double trace_dispatch(const typename RecOf<T>::type *rec, const SegHot *hot,
                                        const SegAux *aux, long long b, int n,
                                        const LineConst &lc, double dnu, double velo,
                                        double istar, double I0, T sf, int &ovs, int &maser);

template <>
__device__ inline double trace_dispatch<double>(const NodeRec<double> *rec, const SegHot *,
                                                const SegAux *, long long b, int n,
                                                const LineConst &lc, double dnu, double velo,
                                                double istar, double I0, double sf, int &ovs,
                                                int &maser) {
    LineK<double> L;
    L.nu0 = lc.nu0; L.c_aa = lc.c_aa; L.hknu = lc.hknu;
    L.A = lc.A; L.Bud = lc.Bud; L.Bdu = lc.Bdu;
    L.dnu = dnu; L.uch = velo;
    return trace_one<double>(rec + b, n, L, I0, istar, sf, ovs, maser);
}

template <>
__device__ inline double trace_dispatch<float>(const NodeLine *rec, const SegHot *hot,
                                               const SegAux *aux, long long b, int n,
                                               const LineConst &, double, double velo,
                                               double istar, double I0, float sf, int &ovs,
                                               int &maser) {
    return (double)trace_split(hot + b, aux + b, rec + b, n, (float)velo, (float)I0,
                               (float)istar, sf, ovs, maser);
}

// image placeholder of a channel left out by NONREDUNDANT skipping (a NaN
// payload no arithmetic produces), replaced by the lazy-continuum pass
constexpr long long SKIPPED_BITS = 0x7ff8dead00000000ll;

// imcir_cmask (telescope.F:546-564): 1 for channel 1, the central beam and
// channels inside the ray's window; the channel integrated only to seed
// the lazy continuum, and its copies, stay 0
__device__ inline unsigned char window_mask(int ray, int ch, const RayInfo *__restrict__ rinfo,
                                            const double *__restrict__ velo, double aks2,
                                            int skip_mode) {
    if (skip_mode != 0 || ray == 0 || ch == 0) return 1;
    const RayInfo ri = rinfo[ray];
    return velo[ch] <= (double)ri.mx + aks2 && velo[ch] >= (double)ri.mn - aks2;
}

// one (ray, channel) of make_image_circular's loop (telescope.F:550-613):
// integrated, or marked for the lazy-continuum pass when NONREDUNDANT
// skipping leaves it out of the ray's window
template <typename T>
__device__ inline double formal_one(int ray, int ch, const long long *__restrict__ off,
                                    const typename RecOf<T>::type *__restrict__ rec,
                                    const SegHot *__restrict__ hot,
                                    const SegAux *__restrict__ aux,
                                    const RayInfo *__restrict__ rinfo,
                                    const double *__restrict__ dnu,
                                    const double *__restrict__ velo,
                                    const double *__restrict__ istar,
                                    const double *__restrict__ i0, const LineConst &lc,
                                    double aks2, int skip_mode, int &ovs, int &maser) {
    const RayInfo ri = rinfo[ray];
    bool doit = true;
    if (skip_mode == 0 && ray > 0 && ch > 0 && ch != ri.cont) {
        doit = velo[ch] <= (double)ri.mx + aks2 && velo[ch] >= (double)ri.mn - aks2;
    }
    if (!doit) return __longlong_as_double(SKIPPED_BITS);   // filled by k_fixup
    T sf = ray == 0 ? (T)lc.starfract : T(0);
    return trace_dispatch<T>(rec, hot, aux, off[ray], ri.n, lc, dnu[ch], velo[ch], istar[ch],
                             i0[ch], sf, ovs, maser);
}

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
// rectangular images (position-velocity cubes)
// ---------------------------------------------------------------------
// a ray of the rectangular imager whose impact parameter is within R*
// takes the stellar intensity at every node that follows a node of the
// innermost radial cell and is itself in it (charintline with rbeam0 = 0,
// telescope.F:4194-4208); the radial index of a node is that of its first
// stencil cell
__global__ void k_star_flags(int nlisted, const int *__restrict__ rays,
                             const long long *__restrict__ off, GeoNode *__restrict__ nodes,
                             int nh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlisted) return;
    const int ray = rays[i];
    const long long b = off[ray], e = off[ray + 1];
    int ir_prev = b < e ? nodes[b].c00 / nh + 1 : 0;
    for (long long k = b + 1; k < e; k++) {
        const int ir = nodes[k].c00 / nh + 1;
        if (ir == 1 && ir_prev == 1) nodes[k].flag |= NF_STAR << 2;
        ir_prev = ir;
    }
}

// one lane per (ray, channel): the intensity and char_tau of
// make_freq_image_rectang (telescope.F:2117-2146).  Pixel rays are ix-major
// (ray = ix*ny + iy); the cube is written iy-major, [iy][ix][channel], and
// rays past the nx*ny pixels (the added star's) into extra[channel]
template <typename T>
__global__ void __launch_bounds__(256)
k_formal_posvel(int nrays, int nfr, int nx, int ny, const long long *__restrict__ off,
                const typename PvRecOf<T>::type *__restrict__ rec,
                const double *__restrict__ dnu, const double *__restrict__ velo,
                const double *__restrict__ istar, const double *__restrict__ i0, LineConst lc,
                double *__restrict__ image, double *__restrict__ tau, double *__restrict__ extra,
                unsigned long long *ovs_out, int *maser_out) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int ovs = 0, maser = 0;
    if (g < (long long)nrays * nfr) {
        const int ray = (int)(g / nfr), ch = (int)(g - (long long)ray * nfr);
        const long long b = off[ray];
        const int n = (int)(off[ray + 1] - b);
        double I = 0.0, t = 0.0;
        if (n > 0) {
            if constexpr (std::is_same<T, double>::value) {
                LineK<double> L;
                L.nu0 = lc.nu0; L.c_aa = lc.c_aa; L.hknu = lc.hknu;
                L.A = lc.A; L.Bud = lc.Bud; L.Bdu = lc.Bdu;
                L.dnu = dnu[ch]; L.uch = velo[ch];
                I = trace_one<double>(rec + b, n, L, i0[ch], istar[ch], 0.0, ovs, maser, &t);
            } else {
                FastLine L{(float)lc.nu0, (float)dnu[ch], (float)velo[ch]};
                float tf = 0.0f;
                I = (double)trace_fast<float>(rec + b, n, L, (float)i0[ch], (float)istar[ch],
                                              0.0f, ovs, maser, &tf);
                t = (double)tf;
            }
        }
        const int npix = nx * ny;
        if (ray < npix) {
            const int ix = ray / ny, iy = ray - ix * ny;
            const size_t o = ((size_t)iy * nx + ix) * nfr + ch;
            image[o] = I;
            tau[o] = t;
        } else {
            extra[(size_t)(ray - npix) * nfr + ch] = I;
        }
    }
    // wave totals of the warnings
    for (int m = 32; m >= 1; m >>= 1) {
        ovs += __shfl_xor(ovs, m);
        maser |= __shfl_xor(maser, m);
    }
    if ((threadIdx.x & 63) == 0) {
        if (ovs) atomicAdd(ovs_out, (unsigned long long)ovs);
        if (maser) atomicOr(maser_out, 1);
    }
}

void launch_star_flags(hipStream_t st, int nlisted, const int *rays, const long long *off,
                       GeoNode *nodes, int nh) {
    if (nlisted <= 0) return;
    hipLaunchKernelGGL(k_star_flags, dim3((nlisted + 63) / 64), dim3(64), 0, st, nlisted, rays,
                       off, nodes, nh);
}

template <typename T>
void launch_posvel(hipStream_t st, const PosvelLaunch &a) {
    auto *rec = reinterpret_cast<typename PvRecOf<T>::type *>(a.rec);
    if constexpr (std::is_same<T, double>::value)
        hipLaunchKernelGGL(k_nodes, dim3(a.nrays), dim3(64), 0, st, a.nrays, a.off, a.geo,
                           a.cells, rec, a.rinfo, a.nfr, a.velo, a.aks2, 1);
    else
        hipLaunchKernelGGL(k_nodes_fast<float>, dim3(a.nrays), dim3(64), 0, st, a.nrays, a.off,
                           a.geo, a.cells, rec, 0ll, a.rinfo, a.nfr, a.velo, a.aks2, 1, a.lc);
    const long long total = (long long)a.nrays * a.nfr;
    hipLaunchKernelGGL(k_formal_posvel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       st, a.nrays, a.nfr, a.nx, a.ny, a.off,
                       reinterpret_cast<const typename PvRecOf<T>::type *>(a.rec), a.dnu, a.velo,
                       a.istar, a.i0, a.lc, a.image, a.tau, a.extra, a.ovs, a.maser);
}
template void launch_posvel<double>(hipStream_t, const PosvelLaunch &);
template void launch_posvel<float>(hipStream_t, const PosvelLaunch &);

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
void launch_cross(hipStream_t st, const GridDev &gd, const double *tan2, const CrossSets &sets,
                  int nmax, int *err) {
    GridView g{gd.nr, gd.ny, gd.rc, gd.tc, gd.rstar, gd.t_sorted};
    int most = 0;
    for (int i = 0; i < sets.n; i++) most = max(most, sets.s[i].nrays);
    dim3 blk(CROSS_BLOCK), grd((most + CROSS_BLOCK - 1) / CROSS_BLOCK, sets.n);
    const size_t lds = (size_t)(gd.nr + 4 + gd.ny + 4 + gd.ny / 2 + 1) * sizeof(double);
    const int in_lds = lds <= 48 * 1024 ? 1 : 0;
    hipLaunchKernelGGL(k_cross, grd, blk, in_lds ? lds : 0, st, g, tan2, sets, nmax, err, in_lds);
}

// velocity [cell][3] and width [cell] into one record per cell; the rays'
// x0 and z0 into one record per ray
__global__ void k_pack_vlw(int ncell, const double *__restrict__ vel,
                           const double *__restrict__ lw, VelLw *__restrict__ out, int nrays,
                           const double *__restrict__ x0, const double *__restrict__ z0,
                           RayXZ *__restrict__ xz) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) out[c] = VelLw{vel[3 * c], vel[3 * c + 1], vel[3 * c + 2], lw[c]};
    if (c < nrays) xz[c] = RayXZ{x0[c], z0[c]};
}

void launch_geo(hipStream_t st, const GridDev &gd, const double *vel, int nh, int nrays,
                int nmax, const long long *off, const double *x0, const double *z0,
                double costh0, double sinth0, const CrossRec *cr,
                GeoNode *nodes, int *err, long long nnodes, const double *lw, SegHot *hot,
                SegAux *aux, VelLw *vlw, RayXZ *xz) {
    GridView g{gd.nr, gd.ny, gd.rc, gd.tc, gd.rstar, gd.t_sorted};
    CellView cv{nh, vel, vlw};
    const size_t tab = (size_t)(gd.nr + 4 + gd.ny + 4) * sizeof(double);
    if (hot) {
        const int ncell = gd.nr * nh, npk = max(ncell, nrays);
        hipLaunchKernelGGL(k_pack_vlw, dim3((npk + 255) / 256), dim3(256), 0, st, ncell, vel,
                           lw, vlw, nrays, x0, z0, xz);
        const dim3 grd((unsigned)((nnodes + 1 + 254) / 255));
        if (tab <= GEO_GRID_LDS_MAX)
            hipLaunchKernelGGL(k_geo_seg<true>, grd, dim3(256), tab, st, g, cv, nmax, nrays,
                               nnodes, off, xz, costh0, sinth0, cr, nodes, hot, aux, err);
        else
            hipLaunchKernelGGL(k_geo_seg<false>, grd, dim3(256), 0, st, g, cv, nmax, nrays,
                               nnodes, off, xz, costh0, sinth0, cr, nodes, hot, aux, err);
    } else {
        const dim3 grd(nrays, (nmax + 255) / 256);
        if (tab <= GEO_GRID_LDS_MAX)
            hipLaunchKernelGGL(k_geo<true>, grd, dim3(256), tab, st, g, cv, nmax, off, x0, z0,
                               costh0, sinth0, cr, nodes, err);
        else
            hipLaunchKernelGGL(k_geo<false>, grd, dim3(256), 0, st, g, cv, nmax, off, x0, z0,
                               costh0, sinth0, cr, nodes, err);
    }
}

void launch_cells(hipStream_t st, int ncell, int nlev, int nlines, int line, int lev_up,
                  int lev_low, double molpg, const double *rho, const double *abund,
                  const double *lw, const double *pops, const double *jd, const double *ad,
                  CellRec *out) {
    hipLaunchKernelGGL(k_cells, dim3((ncell + 255) / 256), dim3(256), 0, st, ncell, nlev,
                       nlines, line, lev_up, lev_low, molpg, rho, abund, lw, pops, jd, ad,
                       out);
}

// ---------------------------------------------------------------------
// batched items: one launch per stage for many (line, inclination)
// items with uneven channel counts
// ---------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(64)
k_nodes_batch(const ItemDev *__restrict__ items, const CellRec *__restrict__ cells_all,
              int ncell, void *rec_all, RayInfo *__restrict__ rinfo_all,
              const double *__restrict__ chan_all) {
    const ItemDev &it = items[blockIdx.y];
    const int ray = blockIdx.x;
    if (ray >= it.nrays) return;
    const CellRec *cells = cells_all + (size_t)it.cslot * ncell;
    const double *velo = chan_all + it.chan0 + it.nfr;
    if constexpr (std::is_same<T, double>::value)
        nodes_ray_exact(ray, threadIdx.x, it.off, it.geo, cells,
                        reinterpret_cast<NodeRec<double> *>(rec_all) + it.rec0,
                        rinfo_all + it.rinfo0, it.nfr, velo, it.aks2, it.skip_mode);
    else
        nodes_ray_line(ray, threadIdx.x, it.off, it.geo, cells,
                       reinterpret_cast<NodeLine *>(rec_all) + it.rec0, rinfo_all + it.rinfo0,
                       it.nfr, velo, it.aks2, it.skip_mode, it.lc);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// segment records of an item: a byte offset from the batch's record buffer
template <typename R>
__device__ inline const R *seg_ptr(const void *rec_all, long long off) {
    return reinterpret_cast<const R *>(static_cast<const char *>(rec_all) + off);
}

// Lane -> (ray, channel) of a tile.  cw == 0: the item's (ray, channel)
// pairs in ray-major order, 64 per tile (a wave may span several rays);
// cw in {16, 32, 64}: 64/cw rays per wave, cw consecutive channels each.
__device__ inline bool tile_lane(int cw, long long lt, int lane, int nrays, int nfr, int &ray,
                                 int &ch) {
    if (cw == 0) {
        const long long g = lt * 64 + lane;
        if (g >= (long long)nrays * nfr) return false;
        ray = (int)(g / nfr);
        ch = (int)(g - (long long)ray * nfr);
        return true;
    }
    const int chunks = (nfr + cw - 1) / cw, rows = 64 / cw;
    const long long rb = lt / chunks;
    const int cc = (int)(lt - rb * chunks);
    ray = (int)(rb * rows) + lane / cw;
    ch = cc * cw + (lane % cw);
    return ray < nrays && ch < nfr;
}

// Persistent formal solution over all tiles of all items: each wave takes
// the next tile from a device counter until every tile is taken, so the
// grid drains by itself; items of any size share the machine.
#define RL_FORMAL_BATCH_PARAMS                                                              \
    const ItemDev *__restrict__ items, int nitems, const long long *__restrict__ tile0,     \
        long long ntiles, int cw, const void *rec_all,                                      \
        const RayInfo *__restrict__ rinfo_all, const double *__restrict__ chan_all,         \
        double *__restrict__ image_all, unsigned long long *__restrict__ ovs_all,            \
        int *__restrict__ maser_all, unsigned char *__restrict__ cmask_all,                 \
        unsigned long long *counter
#define RL_FORMAL_BATCH_ARGS                                                                \
    items, nitems, tile0, ntiles, cw, rec_all, rinfo_all, chan_all, image_all, ovs_all,     \
        maser_all, cmask_all, counter

template <typename T>
__device__ __forceinline__ void formal_batch_body(RL_FORMAL_BATCH_PARAMS) {
    const auto *rec_base = reinterpret_cast<const typename RecOf<T>::type *>(rec_all);
    const int lane = threadIdx.x & 63;
    for (;;) {
        // Every lane executes the atomic (lane 0 adds 1, the others 0) and the
        // tile index is lane 0's result through readfirstlane, so it is
        // wave-uniform (an SGPR) and the loop exit is a scalar branch.  A
        // lane-0-only atomic broadcast with a shuffle leaves the exit
        // condition divergent to the compiler, which may then iterate the
        // other lanes without lane 0.
        const unsigned long long mine = atomicAdd(counter, lane == 0 ? 1ull : 0ull);
        const unsigned lo32 = __builtin_amdgcn_readfirstlane((unsigned)(mine & 0xffffffffull));
        const unsigned hi32 = __builtin_amdgcn_readfirstlane((unsigned)(mine >> 32));
        const unsigned long long t = ((unsigned long long)hi32 << 32) | lo32;
        if (t >= (unsigned long long)ntiles) break;
        // item of the tile: last item with tile0[item] <= t
        int lo = 0, hi = nitems;
        while (hi - lo > 1) {
            const int m = (lo + hi) >> 1;
            if (tile0[m] <= (long long)t) lo = m; else hi = m;
        }
        const ItemDev &it = items[lo];
        int ray = 0, ch = 0, ovs = 0, maser = 0;
        if (tile_lane(cw, (long long)t - tile0[lo], lane, it.nrays, it.nfr, ray, ch)) {
            const double *dnu = chan_all + it.chan0;
            const long long ix = it.img0 + (long long)ray * it.nfr + ch;
            const double v =
                formal_one<T>(ray, ch, it.off, rec_base + it.rec0,
                              seg_ptr<SegHot>(rec_all, it.hot_off),
                              seg_ptr<SegAux>(rec_all, it.aux_off),
                              rinfo_all + it.rinfo0, dnu,
                              dnu + it.nfr, dnu + 2 * it.nfr, dnu + 3 * it.nfr, it.lc, it.aks2,
                              it.skip_mode, ovs, maser);
            image_all[ix] = v;
            if (cmask_all)
                cmask_all[ix] = window_mask(ray, ch, rinfo_all + it.rinfo0, dnu + it.nfr,
                                            it.aks2, it.skip_mode);
        }
        for (int m = 32; m >= 1; m >>= 1) {
            ovs += __shfl_xor(ovs, m);
            maser |= __shfl_xor(maser, m);
        }
        if (lane == 0) {
            if (ovs) atomicAdd(ovs_all + lo, (unsigned long long)ovs);
            if (maser) atomicOr(maser_all + lo, 1);
        }
    }
}

// FP64 (the primary template; FP32 is specialised below): held to 128
// VGPRs, four waves per SIMD instead of three (139 registers)
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
k_formal_batch(RL_FORMAL_BATCH_PARAMS) {
    formal_batch_body<T>(RL_FORMAL_BATCH_ARGS);
}
// FP32 (split records): held to 64 VGPRs, eight waves per SIMD (79
// unconstrained, six waves: 61.4 ms on the 400x200 disk's two
// inclinations; 72, seven waves: 54.1 ms; 64: 52.3 ms).  What spills is
// per-tile state and the sub-grid path's, outside the fast segment loop.
template <>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_formal_batch<float>(RL_FORMAL_BATCH_PARAMS) {
    formal_batch_body<float>(RL_FORMAL_BATCH_ARGS);
}

// The same tiles on a static grid: wave w of block b takes tile 4b + w
// (the hardware dispatcher balances blocks); no tile counter and no tile
// loop.
template <typename T>
__global__ void __launch_bounds__(256)
k_formal_tiles(const ItemDev *__restrict__ items, int nitems,
               const long long *__restrict__ tile0, long long ntiles, int cw,
               const void *rec_all, const RayInfo *__restrict__ rinfo_all,
               const double *__restrict__ chan_all, double *__restrict__ image_all,
               unsigned long long *__restrict__ ovs_all, int *__restrict__ maser_all,
               unsigned char *__restrict__ cmask_all) {
    const auto *rec_base = reinterpret_cast<const typename RecOf<T>::type *>(rec_all);
    const int lane = threadIdx.x & 63;
    const long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;                       // whole waves
    int lo = 0, hi = nitems;
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (tile0[m] <= t) lo = m; else hi = m;
    }
    const ItemDev &it = items[lo];
    int ray = 0, ch = 0, ovs = 0, maser = 0;
    if (tile_lane(cw, t - tile0[lo], lane, it.nrays, it.nfr, ray, ch)) {
        const double *dnu = chan_all + it.chan0;
        const long long ix = it.img0 + (long long)ray * it.nfr + ch;
        const double v =
            formal_one<T>(ray, ch, it.off, rec_base + it.rec0,
                          seg_ptr<SegHot>(rec_all, it.hot_off),
                          seg_ptr<SegAux>(rec_all, it.aux_off),
                          rinfo_all + it.rinfo0, dnu,
                          dnu + it.nfr, dnu + 2 * it.nfr, dnu + 3 * it.nfr, it.lc, it.aks2,
                          it.skip_mode, ovs, maser);
        image_all[ix] = v;
        if (cmask_all)
            cmask_all[ix] = window_mask(ray, ch, rinfo_all + it.rinfo0, dnu + it.nfr, it.aks2,
                                        it.skip_mode);
    }
    for (int m = 32; m >= 1; m >>= 1) {
        ovs += __shfl_xor(ovs, m);
        maser |= __shfl_xor(maser, m);
    }
    if (lane == 0) {
        if (ovs) atomicAdd(ovs_all + lo, (unsigned long long)ovs);
        if (maser) atomicOr(maser_all + lo, 1);
    }
}

template <typename T>
__global__ void k_fixup_batch(const ItemDev *__restrict__ items, const void *rec_all,
                              const RayInfo *__restrict__ rinfo_all,
                              const double *__restrict__ chan_all, double *__restrict__ image_all,
                              unsigned long long *__restrict__ ovs_all,
                              int *__restrict__ maser_all, unsigned char *__restrict__ cmask_all) {
    const int item = blockIdx.y;
    const ItemDev &it = items[item];
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (it.skip_mode != 0 || ray == 0 || ray >= it.nrays) return;
    const RayInfo ri = rinfo_all[it.rinfo0 + ray];
    if (ri.cont < 0) return;
    const auto *rec = reinterpret_cast<const typename RecOf<T>::type *>(rec_all) + it.rec0;
    const double *dnu = chan_all + it.chan0, *velo = dnu + it.nfr, *istar = dnu + 2 * it.nfr,
                 *i0 = dnu + 3 * it.nfr;
    double *row = image_all + it.img0 + (size_t)ray * it.nfr;
    double cont = row[ri.cont];
    int ovs = 0, maser = 0;
    for (int ch = ri.cont + 1; ch < it.nfr; ch++) {
        bool inwin = velo[ch] <= (double)ri.mx + it.aks2 && velo[ch] >= (double)ri.mn - it.aks2;
        if (inwin) continue;
        if (cont != 0.0) {
            row[ch] = cont;
        } else {
            row[ch] = trace_dispatch<T>(rec, seg_ptr<SegHot>(rec_all, it.hot_off),
                                        seg_ptr<SegAux>(rec_all, it.aux_off), it.off[ray], ri.n,
                                        it.lc, dnu[ch],
                                        velo[ch], istar[ch], i0[ch], T(0), ovs, maser);
            cont = row[ch];
        }
    }
    if (ovs) atomicAdd(ovs_all + item, (unsigned long long)ovs);
    if (maser) atomicOr(maser_all + item, 1);
}

__global__ void k_flux_partial_batch(const ItemDev *__restrict__ items,
                                     const double *__restrict__ image_all,
                                     double *__restrict__ partial_all) {
    const ItemDev &it = items[blockIdx.z];
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    const int ir = blockIdx.y;
    if (ch >= it.nfr || ir >= it.nrings) return;
    const double *image = image_all + it.img0;
    double acc = 0.0;
    if (ir == 0) acc = it.center_w * image[ch];
    const double *row = image + (size_t)(1 + ir * it.nphi) * it.nfr + ch;
    double s = 0.0;
    for (int ip = 0; ip < it.nphi; ip++) s = s + row[(size_t)ip * it.nfr];
    s = s / (1.0 * it.nphi);
    acc = acc + s * it.ring_w[ir];
    partial_all[it.part0 + (size_t)ir * it.nfr + ch] = acc;
}

__global__ void k_flux_final_batch(const ItemDev *__restrict__ items,
                                   const double *__restrict__ partial_all, double d2,
                                   double *__restrict__ flux_all) {
    const ItemDev &it = items[blockIdx.y];
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch >= it.nfr) return;
    double s = 0.0;
    for (int k = 0; k < it.nrings; k++) s = s + partial_all[it.part0 + (size_t)k * it.nfr + ch];
    flux_all[it.flux0 + ch] = s / d2;
}

void launch_batch_nodes(hipStream_t st, int fp64, const BatchLaunch &b) {
    dim3 g(b.max_rays, b.nitems);
    if (fp64)
        hipLaunchKernelGGL(k_nodes_batch<double>, g, dim3(64), 0, st, b.items, b.cells, b.ncell,
                           b.rec, b.rinfo, b.chan);
    else
        hipLaunchKernelGGL(k_nodes_batch<float>, g, dim3(64), 0, st, b.items, b.cells, b.ncell,
                           b.rec, b.rinfo, b.chan);
}

void launch_batch_formal(hipStream_t st, int fp64, const BatchLaunch &b, int cw, int nblocks,
                         long long ntiles) {
    if (nblocks == 0) {           // static grid
        const unsigned g = (unsigned)((ntiles + 3) / 4);
        if (fp64)
            hipLaunchKernelGGL(k_formal_tiles<double>, dim3(g), dim3(256), 0, st, b.items,
                               b.nitems, b.tile0, ntiles, cw, b.rec, b.rinfo, b.chan, b.image,
                               b.ovs, b.maser, b.cmask);
        else
            hipLaunchKernelGGL(k_formal_tiles<float>, dim3(g), dim3(256), 0, st, b.items,
                               b.nitems, b.tile0, ntiles, cw, b.rec, b.rinfo, b.chan, b.image,
                               b.ovs, b.maser, b.cmask);
        return;
    }
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, b.counter, 1);
    if (fp64)
        hipLaunchKernelGGL(k_formal_batch<double>, dim3(nblocks), dim3(256), 0, st, b.items,
                           b.nitems, b.tile0, ntiles, cw, b.rec, b.rinfo, b.chan, b.image,
                           b.ovs, b.maser, b.cmask, b.counter);
    else
        hipLaunchKernelGGL(k_formal_batch<float>, dim3(nblocks), dim3(256), 0, st, b.items,
                           b.nitems, b.tile0, ntiles, cw, b.rec, b.rinfo, b.chan, b.image,
                           b.ovs, b.maser, b.cmask, b.counter);
}

void launch_batch_finish(hipStream_t st, int fp64, const BatchLaunch &b) {
    if (b.any_skip) {
        dim3 g((b.max_rays + 127) / 128, b.nitems);
        if (fp64)
            hipLaunchKernelGGL(k_fixup_batch<double>, g, dim3(128), 0, st, b.items, b.rec,
                               b.rinfo, b.chan, b.image, b.ovs, b.maser, b.cmask);
        else
            hipLaunchKernelGGL(k_fixup_batch<float>, g, dim3(128), 0, st, b.items, b.rec,
                               b.rinfo, b.chan, b.image, b.ovs, b.maser, b.cmask);
    }
}

void launch_batch_flux(hipStream_t st, const BatchLaunch &b) {
    dim3 gp((b.max_nfr + 63) / 64, b.max_rings, b.nitems);
    hipLaunchKernelGGL(k_flux_partial_batch, gp, dim3(64), 0, st, b.items, b.image, b.partial);
    dim3 gf((b.max_nfr + 63) / 64, b.nitems);
    hipLaunchKernelGGL(k_flux_final_batch, gf, dim3(64), 0, st, b.items, b.partial, b.d2,
                       b.flux);
}

// The ray positions of a circular image (setup_rays_circular,
// telescope.F:913-1075): ray 0 the central one at the origin, then ring
// ir = 1..nrings, direction k = 0..nphi-1 at x0 = r_ir xh_k, z0 = r_ir zh_k.
// tab: the ring radii [nrings + 1] (entry 0 unused), xh [nphi], zh [nphi],
// formed on the host (their trigonometry needs the host's libm for the
// reference's bits; each product is one IEEE multiply on either side).
__global__ void k_rays_circular(int nrings, int nphi, const double *__restrict__ tab,
                                double *__restrict__ x0, double *__restrict__ z0) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 1 + (long long)nrings * nphi) return;
    if (i == 0) {
        x0[0] = 0.0;
        z0[0] = 0.0;
        return;
    }
    const int ir = (int)((i - 1) / nphi) + 1;
    const int k = (int)(i - 1 - (long long)(ir - 1) * nphi);
    const double r = tab[ir];
    x0[i] = r * tab[nrings + 1 + k];
    z0[i] = r * tab[nrings + 1 + nphi + k];
}

void launch_rays_circular(hipStream_t st, int nrings, int nphi, const double *tab, double *x0,
                          double *z0) {
    const long long n = 1 + (long long)nrings * nphi;
    hipLaunchKernelGGL(k_rays_circular, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       nrings, nphi, tab, x0, z0);
}

void launch_zero_u64(hipStream_t st, unsigned long long *p, int n) {
    if (n > 0) hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
}

int formal_batch_blocks_per_cu(int fp64) {
    int nb = 0;
    if (fp64)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_formal_batch<double>, 256, 0);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_formal_batch<float>, 256, 0);
    return nb > 0 ? nb : 1;
}

// ---------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------
// forward pass of the adjoint: the formal solution on scaled records, one
// lane per (ray, channel), keeping I before every kblk-th segment.  A call
// covers rays ray0.. of the ray set (a chunk): off is the chunk's slice of
// the node CSR (global node indices), rec, ckoff, ckpt and image are the
// chunk's own, node0 its first node.
template <typename T>
__global__ void __launch_bounds__(256)
k_adj_fwd(int nrays, int nfr, int ray0, long long node0, const long long *__restrict__ off,
          const long long *__restrict__ ckoff, const FastRecT<T> *__restrict__ rec,
          const double *__restrict__ dnu, const double *__restrict__ velo,
          const double *__restrict__ istar, const double *__restrict__ i0, LineConst lc,
          double *__restrict__ image, T *__restrict__ ckpt, int kblk) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)nrays * nfr) return;
    const int ray = (int)(g / nfr), ch = (int)(g - (long long)ray * nfr);
    const long long b = off[ray] - node0;
    const int n = (int)(off[ray + 1] - off[ray]);
    if (n < 2) { image[g] = i0[ch]; return; }
    const FastLineT<T> L{(T)lc.nu0, (T)dnu[ch], (T)velo[ch]};
    const T sf = ray0 + ray == 0 ? (T)lc.starfract : T(0);
    const T ist = (T)istar[ch];
    T I = (T)i0[ch];
    FastCarryT<T> c;
    c.u0 = rec[b].u;
    int ovs = 0;
    // I before segments 1, 1 + kblk, 1 + 2 kblk, ... into consecutive slots
    // (a countdown: a remainder by the run-time kblk costs an emulated
    // division per segment)
    T *ck = ckpt + ckoff[ray] * nfr + ch;
    int until = 0;                           // segments until the next checkpoint
    for (int k = 1; k < n; k++) {
        if (until == 0) {
            *ck = I;
            ck += nfr;
            until = kblk;
        }
        until--;
        const FastRecT<T> r = rec[b + k];
        segment_fast<T>(r, rec + b + k - 1, point_of<T>(r), L, I, ist, sf, c, ovs);
    }
    image[g] = (double)I;
}

// Node gradients of a block, summed over its lanes in batches of RED_N
// nodes through LDS.  A block holds one ray and one wave per channel group
// (all waves visit the same nodes in the same order).  Each lane writes its
// values in row `lane` of its wave's [64][RED_W] area (RED_W words, odd, so
// that the writes of 64 lanes fall in distinct banks); lane p of each wave
// sums column p (node p/5, quantity p%5) over its 64 rows into the wave's
// partial sums; wave 0 then adds the waves' partial sums in wave order to
// gnode.  The block owns its ray's node slots and every sum runs in a fixed
// order: the gradients repeat bit for bit.  Seven nodes: 64 x 37 x 4 B =
// 9.3 KB per wave, so that the sixteen waves per CU that the reverse
// sweep's registers allow also fit in LDS (eight nodes, 10.5 KB, left room
// for fifteen: 13.7 -> 12.6 ms on the 200x100 disk).
constexpr int RED_N = 7;
constexpr int RED_W = RED_N * 5 + 2;

template <typename T>
__host__ __device__ inline size_t adj_rev_lds(int nw) {
    size_t b = (size_t)nw * 64 * RED_W * sizeof(T);
    b += ((size_t)nw * RED_N * 5 * sizeof(T) + 7) & ~(size_t)7;
    return b + RED_N * sizeof(long long);
}

template <typename T>
struct BlockNodeSum {
    T (*row)[RED_W];           // this wave's [64][RED_W]
    T *part;                   // [nw][RED_N * 5] the waves' column sums
    long long *node;           // [RED_N]
    int nw, w;                 // waves in the block, this wave
    int n = 0;

    __device__ void flush(double *gnode) {
        __syncthreads();
        const int lane = threadIdx.x & 63;
        for (int p = lane; p < n * 5; p += 64) {
            T acc = T(0);
            for (int i = 0; i < 64; i++) acc += row[i][p];
            part[w * (RED_N * 5) + p] = acc;
        }
        __syncthreads();
        if (w == 0) {
            for (int p = lane; p < n * 5; p += 64) {
                T acc = part[p];
                for (int v = 1; v < nw; v++) acc += part[v * (RED_N * 5) + p];
                if (acc != T(0)) {
                    double *g = gnode + node[p / 5] * 5 + (p % 5);
                    *g = *g + (double)acc;
                }
            }
        }
        n = 0;
    }
    __device__ void push(double *gnode, long long nd, T a0, T a1, T a2, T a3, T a4) {
        T *r = row[threadIdx.x & 63] + n * 5;
        r[0] = a0; r[1] = a1; r[2] = a2; r[3] = a3; r[4] = a4;
        if (threadIdx.x == 0) node[n] = nd;
        if (++n == RED_N) flush(gnode);
    }
};

// line values carried into segment k: node k-1 with segment k-1's profile
// (z, phi: its argument and value, the end-point profile of segment k-1)
template <typename T>
__device__ inline void carried_in(const FastRecT<T> &a, const FastLineT<T> &L, T &jl, T &al,
                                  T &z, T &phi) {
    profile_at<T>(L, a.u, a.inv_aa, a.norm, z, phi);
    jl = a.jch * phi;
    al = a.dch * phi;
}

// reverse pass: one wave per ray, over its channels in groups of 64 (in
// order); per node the gradients on (jd, ad, jc, dc) / sig and per segment
// on the mean width, summed over the channels into gnode (which the block
// owns for its ray's nodes: no atomics, a fixed summation order)
// (FP32: held to 128 VGPRs, four waves per SIMD instead of three; it was
// 130, and the latency of the per-segment record loads and node sums needs
// the extra wave more than the two registers)
template <typename T>
__device__ __forceinline__ void adj_rev_group(
    int grp, int ray, int nfr, int nphi, int ray0, long long node0,
    const long long *__restrict__ off, const long long *__restrict__ ckoff,
    const FastRecT<T> *__restrict__ rec, const double *__restrict__ dnu,
    const double *__restrict__ velo, const double *__restrict__ istar, const LineConst &lc,
    const double *__restrict__ gflux, const double *__restrict__ ring_w, double center_w,
    double d2, const T *__restrict__ ckpt, double *__restrict__ gnode, int kblk,
    BlockNodeSum<T> &red) {
    const int lane = threadIdx.x & 63;
    const int ch = grp * 64 + lane;
    const bool active = ch < nfr;
    const int chc = active ? ch : nfr - 1;
    const long long b = off[ray] - node0;    // chunk-local node index
    const int n = (int)(off[ray + 1] - off[ray]);
    const FastLineT<T> L{(T)lc.nu0, (T)dnu[chc], (T)velo[chc]};
    // flux = (center_w I_0 + sum_rings ring_w mean_phi I) / d2 (telescope.F:1386-1418)
    const int rg = ray0 + ray;               // index in the ray set
    const double w = rg == 0 ? center_w : ring_w[(rg - 1) / nphi] / (1.0 * nphi);
    T lam = active ? (T)(gflux[ch] * w / d2) : T(0);
    const T sf = rg == 0 ? (T)lc.starfract : T(0);
    const T ist = (T)istar[chc];
    const T cai = (T)lc.c_aa;
    T gjl = T(0), gal = T(0);
    PGrad<T> cur, prv;
    pg_zero(cur);
    pg_zero(prv);
    T isub[33];                    // 33 entries: kept in scratch memory, not registers
    const int nseg = n - 1;
    // profile at the end node of the segment being reversed, with its
    // constants: computed for the last segment, then carried from the
    // previous iteration's carried-in values (the same quantity)
    T zc, phic;
    {
        const FastRecT<T> r = rec[b + nseg];
        profile_at<T>(L, r.u, r.inv_aa, r.norm, zc, phic);
    }
    // reverse of segment k (nodes k-1 -> k) from I at its start
    auto reverse = [&](int k, T Is) {
        const FastRecT<T> r = rec[b + k];
        const FastRecT<T> a = rec[b + k - 1];
        T g_lnaa = T(0), g_dls = T(0);
        T jl0c = T(0), al0c = T(0), za = T(0), phia = T(0);
        // every lane (lanes past the last channel repeat it with lam = 0,
        // so their gradients are 0)
        if (k > 1) carried_in<T>(a, L, jl0c, al0c, za, phia);
        lam = segment_rev<T>(r, point_of<T>(a), point_of<T>(r), L, Is, lam, ist, sf,
                             k == 1 ? 1 : 0, jl0c, al0c, zc, phic, gjl, gal, prv, cur, g_lnaa,
                             g_dls, isub);
        zc = za;
        phic = phia;
        // d lw_mean: phi through inv_aa = 1/(c_aa lw), sub-grid width dls ~ lw
        const T g_lw = (g_dls * r.dls - g_lnaa) * cai * r.inv_aa;
        red.push(gnode, b + k, cur.jd, cur.ad, cur.jc, cur.dc, g_lw);
        cur = prv;
        pg_zero(prv);
    };
    const T *ck = ckpt + ckoff[ray] * nfr + chc;
    if (kblk == 1) {
        // I before every segment was kept: nothing to recompute; the next
        // segment's start value is loaded one iteration ahead
        T Inext = ck[(long long)(nseg - 1) * nfr];
        for (int k = nseg; k >= 1; k--) {
            const T Is = Inext;
            if (k > 1) Inext = ck[(long long)(k - 2) * nfr];
            reverse(k, Is);
        }
    } else {
        T iseg[ADJ_K];
        const int nblk = (nseg + kblk - 1) / kblk;
        for (int blk = nblk - 1; blk >= 0; blk--) {
            const int k0 = blk * kblk + 1, k1 = min(k0 + kblk - 1, nseg);
            // forward over the block from its checkpoint
            T I = ck[(long long)blk * nfr];
            FastCarryT<T> c;
            c.init = k0 == 1 ? 1 : 0;
            if (k0 > 1) {
                const FastRecT<T> a = rec[b + k0 - 1];
                T jl0, al0, z_, phi_;
                carried_in<T>(a, L, jl0, al0, z_, phi_);
                carry_start<T>(point_of<T>(a), jl0, al0, c);
            }
            c.u0 = rec[b + k0 - 1].u;
            int ovs = 0;
            for (int k = k0; k <= k1; k++) {
                iseg[k - k0] = I;
                if (k == k1) break;                // the block's last segment is only reversed
                const FastRecT<T> r = rec[b + k];
                if (active)
                    segment_fast<T>(r, rec + b + k - 1, point_of<T>(r), L, I, ist, sf, c, ovs);
            }
            // backwards over the block
            for (int k = k1; k >= k0; k--) reverse(k, iseg[k - k0]);
        }
    }
    red.push(gnode, b, cur.jd, cur.ad, cur.jc, cur.dc, T(0));
    red.flush(gnode);
}

// one block of nw waves per ray: wave w takes channel groups w, w + nw, ...
// (a group past the last channel runs with every lane idle, so that the
// block's waves meet at the same node batches)
#define RL_ADJ_REV_PARAMS                                                                   \
    int nrays, int nfr, int nphi, int ray0, long long node0,                                \
        const long long *__restrict__ off, const long long *__restrict__ ckoff,             \
        const FastRecT<T> *__restrict__ rec, const double *__restrict__ dnu,                \
        const double *__restrict__ velo, const double *__restrict__ istar, LineConst lc,    \
        const double *__restrict__ gflux, const double *__restrict__ ring_w, double center_w, \
        double d2, const T *__restrict__ ckpt, double *__restrict__ gnode, int kblk, int nw
template <typename T>
__device__ __forceinline__ void adj_rev_body(RL_ADJ_REV_PARAMS) {
    extern __shared__ __align__(16) unsigned char adj_lds[];
    const int ray = blockIdx.x;
    if (off[ray + 1] - off[ray] < 2) return;     // uniform over the block
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    BlockNodeSum<T> red;
    red.row = reinterpret_cast<T (*)[RED_W]>(adj_lds) + (size_t)w * 64;
    red.part = reinterpret_cast<T *>(adj_lds + (size_t)nw * 64 * RED_W * sizeof(T));
    red.node = reinterpret_cast<long long *>(
        adj_lds + adj_rev_lds<T>(nw) - RED_N * sizeof(long long));
    red.nw = nw;
    red.w = w;
    const int ngrp = (nfr + 63) / 64;
    const int npass = (ngrp + nw - 1) / nw;
    for (int p = 0; p < npass; p++)
        adj_rev_group<T>(p * nw + w, ray, nfr, nphi, ray0, node0, off, ckoff, rec, dnu, velo,
                         istar, lc, gflux, ring_w, center_w, d2, ckpt, gnode, kblk, red);
}

// waves per block at most: FP64 2 (its registers allow two waves per SIMD),
// FP32 3.  The block's waves meet at every node batch, each waiting for the
// slowest; on the 1024x512 disk's 300 channels (five groups) an adjoint
// item took 3.97 s with one wave per block (the groups in turn), 3.41 s
// with two, 3.28 s with three and 4.45 s with five.
template <typename T> struct AdjRevWaves { static constexpr int max = 2; };
template <> struct AdjRevWaves<float> { static constexpr int max = 3; };

template <typename T>
__global__ void __launch_bounds__(64 * 2) __attribute__((amdgpu_waves_per_eu(2)))
k_adj_rev(RL_ADJ_REV_PARAMS) {
    adj_rev_body<T>(nrays, nfr, nphi, ray0, node0, off, ckoff, rec, dnu, velo, istar, lc, gflux,
                    ring_w, center_w, d2, ckpt, gnode, kblk, nw);
}
template <>
__global__ void __launch_bounds__(64 * 3) __attribute__((amdgpu_waves_per_eu(4)))
k_adj_rev<float>(int nrays, int nfr, int nphi, int ray0, long long node0,
                 const long long *__restrict__ off, const long long *__restrict__ ckoff,
                 const FastRecT<float> *__restrict__ rec, const double *__restrict__ dnu,
                 const double *__restrict__ velo, const double *__restrict__ istar, LineConst lc,
                 const double *__restrict__ gflux, const double *__restrict__ ring_w,
                 double center_w, double d2, const float *__restrict__ ckpt,
                 double *__restrict__ gnode, int kblk, int nw) {
    adj_rev_body<float>(nrays, nfr, nphi, ray0, node0, off, ckoff, rec, dnu, velo, istar, lc,
                        gflux, ring_w, center_w, d2, ckpt, gnode, kblk, nw);
}

// ---- node -> cell transpose of the interpolation stencil (once per ray set)
__device__ inline void stencil_weights(const GeoNode &gn, double w[4]) {
    const double dr = gn.dr, dt = gn.dt;
    const int icr = gn.flag & 3;
    w[0] = w[1] = w[2] = w[3] = 0.0;
    if (icr == 1) { w[0] = 1.0 - dt; w[1] = dt; }
    else if (icr == 2) { w[0] = 1.0 - dr; w[2] = dr; }
    else {
        w[0] = (1.0 - dr) * (1.0 - dt); w[1] = (1.0 - dr) * dt;
        w[2] = dr * (1.0 - dt); w[3] = dr * dt;
    }
}

// entry 4k+j: stencil cell j of node k; key ncell for a zero weight
__global__ void k_stencil_entries(long long nnodes, const GeoNode *__restrict__ geo, int ncell,
                                  int *__restrict__ keys, int *__restrict__ vals,
                                  double *__restrict__ w) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnodes) return;
    const GeoNode gn = geo[k];
    double wk[4];
    stencil_weights(gn, wk);
    const int c[4] = {gn.c00, gn.c01, gn.c10, gn.c11()};
    for (int j = 0; j < 4; j++) {
        const long long e = 4 * k + j;
        keys[e] = wk[j] != 0.0 ? c[j] : ncell;
        vals[e] = (int)e;
        w[e] = wk[j];
    }
}

// off[c] = first sorted entry of cell c (binary search; off[ncell] = end
// of the entries with a weight)
__global__ void k_cell_offsets(int ncell, long long nent, const int *__restrict__ keys,
                               int *__restrict__ off) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    long long lo = 0, hi = nent;
    while (lo < hi) {
        const long long m = (lo + hi) >> 1;
        if (keys[m] < c) lo = m + 1; else hi = m;
    }
    off[c] = (int)lo;
}

// node gradients in scaled units -> true node values: jd, ad, lw (the
// mean width of segments k and k+1 split to their nodes), N_up, N_low
__global__ void k_node_conv(long long nnodes, const GeoNode *__restrict__ geo,
                            const CellRec *__restrict__ cells, const double *__restrict__ gnode,
                            LineConst lc, double *__restrict__ gtrue) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnodes) return;
    const GeoNode gn = geo[k];
    double v[5], sig, jc, dc;
    node_interp(gn, cells, v);
    scaled_node(v, lc, sig, jc, dc);
    const double *g = gnode + k * 5;
    const double g_jc = g[2] / sig, g_dc = g[3] / sig;
    double *o = gtrue + k * 5;
    o[CR_JD] = g[0] / sig;
    o[CR_AD] = g[1] / sig;
    // a ray's first node carries no segment gradient, so the next node's
    // slot is 0 across rays
    o[CR_LW] = 0.5 * (g[4] + (k + 1 < nnodes ? g[9] : 0.0));
    o[CR_NU] = lc.hknu * lc.A * g_jc - lc.hknu * lc.Bud * g_dc;   // jc = hknu N_u A
    o[CR_NL] = lc.hknu * lc.Bdu * g_dc;                           // dc = hknu (N_l B_lu - N_u B_ul)
}

// one wave per cell: the stencil-weighted sum of its nodes' gradients,
// then N = pop * abund * rho * m_pg (line.F:3998) to the trace's inputs
__global__ void __launch_bounds__(64)
k_cell_gather(int ncell, const int *__restrict__ off, const int *__restrict__ vals,
              const double *__restrict__ w, const double *__restrict__ gtrue, int nlev,
              int nlines, int line, int lev_up, int lev_low, double molpg,
              const double *__restrict__ rho, const double *__restrict__ abund,
              const double *__restrict__ pops, double *__restrict__ g_rho,
              double *__restrict__ g_abund, double *__restrict__ g_lw,
              double *__restrict__ g_pops, double *__restrict__ g_jd,
              double *__restrict__ g_ad) {
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = off[c] + lane; i < off[c + 1]; i += 64) {
        const int e = vals[i];
        const double we = w[e];
        const double *g = gtrue + (long long)(e >> 2) * 5;
        for (int f = 0; f < 5; f++) a[f] += we * g[f];
    }
    for (int m = 32; m >= 1; m >>= 1)
        for (int f = 0; f < 5; f++) a[f] += __shfl_xor(a[f], m);
    if (lane != 0) return;
    const double gnu = a[CR_NU], gnl = a[CR_NL];
    const double pu = pops[(size_t)c * nlev + lev_up], pl = pops[(size_t)c * nlev + lev_low];
    const double ab = abund[c], rh = rho[c];
    g_jd[(size_t)c * nlines + line] += a[CR_JD];
    g_ad[(size_t)c * nlines + line] += a[CR_AD];
    g_lw[c] += a[CR_LW];
    g_pops[(size_t)c * nlev + lev_up] += gnu * ab * rh * molpg;
    g_pops[(size_t)c * nlev + lev_low] += gnl * ab * rh * molpg;
    g_abund[c] += (gnu * pu + gnl * pl) * rh * molpg;
    g_rho[c] += (gnu * pu + gnl * pl) * ab * molpg;
}

int build_stencil_transpose(hipStream_t st, long long nnodes, const GeoNode *geo, int ncell,
                            int *keys_a, int *keys_b, int *vals_a, int *vals_b, double *w,
                            int *off, void *temp, size_t *temp_bytes) {
    const long long nent = 4 * nnodes;
    if (temp == nullptr) {
        return (int)hipcub::DeviceRadixSort::SortPairs(nullptr, *temp_bytes, keys_a, keys_b,
                                                       vals_a, vals_b, (int)nent, 0, 32, st);
    }
    hipLaunchKernelGGL(k_stencil_entries, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0,
                       st, nnodes, geo, ncell, keys_a, vals_a, w);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_a, keys_b, vals_a,
                                                      vals_b, (int)nent, 0, 32, st);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_cell_offsets, dim3((ncell + 1 + 255) / 256), dim3(256), 0, st, ncell,
                       nent, keys_b, off);
    return 0;
}

// node records and the checkpointed forward of one chunk of an item's rays
template <typename T>
void launch_adjoint_fwd(hipStream_t st, const AdjLaunch &a) {
    auto *rec = reinterpret_cast<FastRecT<T> *>(a.rec);
    auto *ck = reinterpret_cast<T *>(a.ckpt);
    hipLaunchKernelGGL(k_nodes_fast<T>, dim3(a.nrays), dim3(64), 0, st, a.nrays, a.off, a.geo,
                       a.cells, rec, a.node0, a.rinfo, a.nfr, a.velo, a.aks2, 1, a.lc);
    const long long total = (long long)a.nrays * a.nfr;
    hipLaunchKernelGGL(k_adj_fwd<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       a.nrays, a.nfr, a.ray0, a.node0, a.off, a.ckoff, rec, a.dnu, a.velo,
                       a.istar, a.i0, a.lc, a.image, ck, a.kblk);
}

// the spectrum of an item from its whole image (every chunk traced)
void launch_adjoint_flux(hipStream_t st, const AdjLaunch &a) {
    dim3 gp((a.nfr + 63) / 64, a.nchunks);
    hipLaunchKernelGGL(k_flux_partial, gp, dim3(64), 0, st, a.nfr, a.nrings, a.nphi, 1, a.image,
                       a.ring_w, a.center_w, a.partial);
    hipLaunchKernelGGL(k_flux_final, dim3((a.nfr + 63) / 64), dim3(64), 0, st, a.nfr,
                       a.nchunks, a.partial, a.d2, a.flux);
}

// reverse sweep of one item (records and checkpoints from
// launch_adjoint_fwd), node gradients to the cells
template <typename T>
void launch_adjoint_rev(hipStream_t st, const AdjLaunch &a) {
    auto *rec = reinterpret_cast<const FastRecT<T> *>(a.rec);
    auto *ck = reinterpret_cast<const T *>(a.ckpt);
    // one wave per channel group, as many as 64 KB of LDS holds
    const int ngrp = (a.nfr + 63) / 64;
    int nw = std::min(ngrp, AdjRevWaves<T>::max);
    while (nw > 1 && adj_rev_lds<T>(nw) > 65536) nw--;
    hipLaunchKernelGGL(k_adj_rev<T>, dim3(a.nrays), dim3(64 * nw), adj_rev_lds<T>(nw), st,
                       a.nrays, a.nfr, a.nphi, a.ray0, a.node0, a.off, a.ckoff, rec, a.dnu,
                       a.velo, a.istar, a.lc, a.gflux, a.ring_w, a.center_w, a.d2, ck, a.gnode,
                       a.kblk, nw);
    hipLaunchKernelGGL(k_node_conv, dim3((unsigned)((a.nnodes + 255) / 256)), dim3(256), 0, st,
                       a.nnodes, a.geo + a.node0, a.cells, a.gnode, a.lc, a.gtrue);
    hipLaunchKernelGGL(k_cell_gather, dim3(a.ncell), dim3(64), 0, st, a.ncell, a.st_off,
                       a.st_vals, a.st_w, a.gtrue, a.nlev, a.nlines, a.line, a.lev_up,
                       a.lev_low, a.molpg, a.rho, a.abund, a.pops, a.g_rho, a.g_abund, a.g_lw,
                       a.g_pops, a.g_jd, a.g_ad);
}

template void launch_adjoint_fwd<double>(hipStream_t, const AdjLaunch &);
template void launch_adjoint_fwd<float>(hipStream_t, const AdjLaunch &);
template void launch_adjoint_rev<double>(hipStream_t, const AdjLaunch &);
template void launch_adjoint_rev<float>(hipStream_t, const AdjLaunch &);

int adjoint_block_len() { return ADJ_K; }
size_t fast_rec_size(int fp64) { return fp64 ? sizeof(FastRecT<double>) : sizeof(FastRecT<float>); }


}  // namespace rl
