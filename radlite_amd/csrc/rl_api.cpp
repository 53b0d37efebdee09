// C ABI of the raytracer (include/radlite.h): context, device tables,
// ray sets, and the per-item launch sequence.  Host arithmetic that
// feeds the kernels (ring radii, passband, line constants) follows the
// reference's evaluation order so that FP64 mode reproduces it.
#include "radlite.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "rl_device.h"
#include "rl_engine.h"

using rl::CellRec;
using rl::GeoNode;
using rl::RayInfo;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            if (p) (void)hipFree(p);
            p = o.p; n = o.n; o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// One item's adjoint state: node records, the checkpoint CSR and the I
// checkpoints.  The tape keeps one per item between
// rl_trace_lines_forward_tape and rl_trace_lines_adjoint.
struct AdjItemBufs {
    DevBuf rec, ckoff, ckpt;
};

struct RaySet {
    double theta0 = 0.0, costh0 = 1.0, sinth0 = 0.0;
    int nrays = 0, nrings = 0, nphi = 0;
    std::vector<double> x0, z0, ring_r, ring_ri;
    std::vector<double> raytab;       // circular: ring radii and directions (k_rays_circular)
    std::vector<double> ring_w;       // flux weights of the rings (uploaded from here)
    DevBuf d_x0, d_z0, d_raytab;
    // trajectory CSR (built lazily, shared by every line of this inclination;
    // the nodes may be evicted to make room and are then built again)
    bool counted = false;             // off / nnodes known
    bool traj_ready = false;          // d_nodes built
    unsigned long long used = 0;      // last use, for eviction
    long long nnodes = 0;
    std::vector<long long> off;
    DevBuf d_off, d_nodes;
    DevBuf d_ringw;
    // rectangular image (position-velocity cube) ray set: pixel rays ix
    // outer, iy inner, then the added-star ray (central, or -1)
    bool rect = false;
    int nx = 0, ny = 0, starunres = 0, central = -1;
    double spx = 0.0, spy = 0.0;
    std::vector<double> bpix;
    DevBuf d_bpix;
    std::vector<int> star_rays;        // traced rays with impact parameter <= R*
    DevBuf d_star_rays;
    // FP32 segment records (rl_fwd32.h), line-independent: built with the
    // nodes, rebuilt when the line widths change (lw_gen)
    bool seg_ready = false;
    unsigned long long seg_lw_gen = 0;
    DevBuf d_hot, d_aux;
    // node -> cell stencil transpose for the adjoint (built on first use)
    bool st_ready = false;
    DevBuf d_st_vals, d_st_w, d_st_off;
    double center_w = 0.0;
};

}  // namespace

struct rl_ctx {
    int device = 0, precision = RL_FP64;
    hipStream_t stream = nullptr;
    std::string err;
    // grid
    int nr = 0, nh = 0;
    std::vector<double> rc, tc;
    int t_sorted = 0;                      // the mirrored theta grid is non-decreasing
    DevBuf d_rc, d_tc, d_tan2;
    // medium
    bool have_medium = false;
    double umass_av = 2.3, aksmax = 0.0;
    DevBuf d_rho, d_abund, d_vel, d_lw;
    DevBuf d_vlw, d_xz;                    // velocity and width per cell, x0 and z0 per ray,
                                           // packed per node pass
    int nlev = 0;
    DevBuf d_pops;
    // lines
    int nlines = 0;
    std::vector<double> nu0, A, gu, gl;
    std::vector<int> lev_up, lev_low;
    DevBuf d_jd, d_ad;
    bool have_dust = false;
    // star, boundary
    double rstar = 0.0;
    std::vector<double> cont_nu, istar;
    int outer_bc = 0;
    std::vector<double> isrf_nu, isrf;     // interstellar field (outer boundary type 3)
    // rays
    std::vector<RaySet> incl;
    // per-item work buffers
    DevBuf d_cells, d_rec, d_rinfo, d_image, d_chan, d_partial, d_flux, d_err, d_ovs,
        d_maser;
    // crossing-pass scratch, one slot per ray set of a launch: its
    // per-ray cone roots and its crossing slots; x_owner = the ray set
    // whose crossings a slot holds (-1: none)
    std::vector<DevBuf> x_scr, x_cross;
    std::vector<int> x_owner;
    int cells_line = -1;
    std::vector<double> h_chan, h_vel;
    // node buffers of the ray sets a setup call replaced, reused by the next
    // builds of the same sizes (counted against the budget; freed first when
    // room is needed)
    std::vector<DevBuf> spare;
    std::vector<DevBuf> x_cnt;               // k_cross node counts per slot
    // batching
    DevBuf d_items, d_tile0, d_counter, d_cmask;
    std::vector<rl::ItemDev> h_items;
    std::vector<long long> h_tile0;
    size_t mem_budget = 0;            // 0: a share of the free device memory
    size_t auto_budget = 0;
    unsigned long long tick = 0;      // use counter of the ray sets
    size_t last_work = 0;             // work bytes of the last call's largest unit
    int last_chunks = 0;
    bool use_graph = true, autotune = true;
    int fixed_cw = -1, fixed_persistent = 1, last_cw = 0, last_blocks = 0, last_batches = 0;
    int last_kblk = 0, fixed_kblk = 0, last_used_tape = 0;
    // inputs generation: bumped by every setter, so that a tape taken
    // before a change is not used after it
    unsigned long long gen = 1;
    // bumped whenever the line-width field changes (the FP32 segment
    // records depend on it)
    unsigned long long lw_gen = 1;
    std::map<std::string, std::pair<int, int>> tune;
    std::map<std::string, hipGraphExec_t> graphs;
    // adjoint buffers; the tape of rl_trace_lines_forward_tape
    std::vector<AdjItemBufs> tape;
    std::vector<rl_item> tape_items;
    std::vector<double> tape_flux;
    unsigned long long tape_gen = 0;
    int tape_kblk = 0;
    DevBuf d_stc_vals, d_stc_w, d_stc_off;    // a split item's chunk transpose
    DevBuf d_arec, d_ckpt, d_ckoff, d_gflux, d_gnode, d_gcell, d_grho, d_gabund, d_glw,
        d_gpops, d_gjd, d_gad;
    // device pre-pass: level data, dust species at the line centres, and the
    // last physical fields (for its adjoint)
    int pre_nlev = 0, pre_nspec = 0, pre_dust_from_gas = 0, pre_ready = 0;
    double pre_q = 0.0, pre_cw = 0.0;
    int pre_alpha_on = 0;                    // the alpha turbulence model (rl_set_turbulence_alpha)
    double pre_alpha = 0.0, pre_gamma = 0.0, pre_mu = 0.0;
    int pre_vt_explicit = 1;                 // the last rl_set_physics took v_turb
    int pre_nq = 0;                          // partition sum table (rl_set_partition_table)
    DevBuf d_qx, d_qy;
    DevBuf d_EK, d_g, d_nu0, d_rhod, d_Td, d_kabs, d_kscat, d_T, d_la, d_vt, d_part;
    DevBuf d_scat, d_scat_lo, d_scat_w;     // scattering source (rl_set_dust_scatter)
    int scat_nf = 0;
    DevBuf d_gT, d_gTd, d_gvt, d_gla;
    // the library's fixed device workspace on this context's device (for
    // rl_trace_lines_forward_noalloc), or the lookup's failure
    void *scratch = nullptr;
    std::string scratch_err;
    // stats
    double t_ms[5] = {0, 0, 0, 0, 0};
    long long counts[4] = {0, 0, 0, 0};
    hipEvent_t ev[6] = {};
};

static int fail(rl_ctx *c, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
static int fail(rl_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, x)                                                              \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess)                                                       \
            return fail(ctx, RL_E_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                        \
    } while (0)

static int upload(rl_ctx *c, DevBuf &b, const void *src, size_t bytes) {
    HIPCHK(c, b.ensure(bytes));
    HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    return 0;
}

// ---------------------------------------------------------------------
// context
// ---------------------------------------------------------------------
extern "C" int rl_create(rl_ctx **out, int device, int precision) {
    if (!out) return RL_E_ARG;
    *out = nullptr;
    if (precision != RL_FP64 && precision != RL_FP32) return RL_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
        return RL_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return RL_E_HIP;
    rl_ctx *c = new rl_ctx();
    c->device = device;
    c->precision = precision;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RL_E_HIP;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) { delete c; return RL_E_HIP; }
    // the fixed workspace's device address, once per device (the entry
    // points that do not use it work without it)
    const hipError_t se = rl::scratch_attach(device, &c->scratch);
    if (se != hipSuccess) c->scratch_err = hipGetErrorString(se);
    *out = c;
    return 0;
}

extern "C" void rl_destroy(rl_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *rl_last_error(const rl_ctx *c) {
    return c ? c->err.c_str() : "no context";
}

extern "C" int rl_sync(rl_ctx *c) {
    if (!c) return RL_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_last_stats(const rl_ctx *c, double *t, long long *n) {
    if (!c) return RL_E_ARG;
    if (t) for (int i = 0; i < 5; i++) t[i] = c->t_ms[i];
    if (n) for (int i = 0; i < 4; i++) n[i] = c->counts[i];
    return 0;
}

// ---------------------------------------------------------------------
// inputs
// ---------------------------------------------------------------------
extern "C" int rl_set_grid(rl_ctx *c, int nr, const double *r, int nh, const double *th) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !r || !th || nr < 3 || nh < 2) return fail(c, RL_E_ARG, "rl_set_grid: bad args");
    HIPCHK(c, hipSetDevice(c->device));
    for (int i = 1; i < nr; i++)
        if (!(r[i] > r[i - 1])) return fail(c, RL_E_ARG, "radius grid not increasing");
    for (int i = 0; i < nh; i++)
        if (th[i] < 0.0 || th[i] > 1.5707963268)          // grid.F:1150-1157
            return fail(c, RL_E_ARG, "theta.inp value out of [0, pi/2]");
    const int ny = 2 * nh;
    c->nr = nr;
    c->nh = nh;
    c->rc.assign(nr + 4, 0.0);
    c->tc.assign(ny + 4, 0.0);
    auto R = [&](int i) -> double & { return c->rc[i + 1]; };
    auto T = [&](int i) -> double & { return c->tc[i + 1]; };
    for (int i = 1; i <= nr; i++) R(i) = r[i - 1];
    R(0) = R(1) * R(1) / R(2);                              // grid.F:753-760
    R(-1) = R(0) * R(0) / R(1);
    R(nr + 1) = R(nr) * R(nr) / R(nr - 1);
    R(nr + 2) = R(nr + 1) * R(nr + 1) / R(nr);
    for (int i = 1; i <= nh; i++) T(i) = th[i - 1];
    for (int i = 1; i <= nh; i++) T(2 * nh + 1 - i) = 3.14159265359 - T(i);  // :1160-1163
    const float twopi_sp = 2 * 3.1415926f;                 // grid.F:1173-1176
    T(0) = -T(1);
    T(-1) = -T(2);
    T(ny + 1) = (double)twopi_sp - T(ny);
    T(ny + 2) = (double)twopi_sp - T(ny - 1);
    c->t_sorted = 1;
    for (int i = 1; i < ny; i++)
        if (!(T(i + 1) >= T(i))) c->t_sorted = 0;
    if (int e = upload(c, c->d_rc, c->rc.data(), c->rc.size() * 8)) return e;
    if (int e = upload(c, c->d_tc, c->tc.data(), c->tc.size() * 8)) return e;
    // tan^2 of the cone angles (telescope.F:2973), ray independent
    std::vector<double> t2(nh + 1, 0.0);
    for (int iy = 1; iy <= nh; iy++) {
        double tn = std::tan(T(iy));
        t2[iy] = tn * tn;
    }
    if (int e = upload(c, c->d_tan2, t2.data(), t2.size() * 8)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->incl.clear();
    c->x_owner.assign(c->x_owner.size(), -1);
    c->cells_line = -1;
    return 0;
}

extern "C" int rl_set_medium(rl_ctx *c, const double *rho, const double *abund,
                             const double *vel, const double *lw, double umass_av,
                             double aksmax) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !c->nr) return fail(c, RL_E_ARG, "rl_set_medium: set the grid first");
    if (!rho || !abund || !vel || !lw || !(umass_av > 0.0))
        return fail(c, RL_E_ARG, "rl_set_medium: bad args");
    HIPCHK(c, hipSetDevice(c->device));
    size_t nc = (size_t)c->nr * c->nh;
    if (int e = upload(c, c->d_rho, rho, nc * 8)) return e;
    if (int e = upload(c, c->d_abund, abund, nc * 8)) return e;
    const bool same_vel = c->h_vel.size() == nc * 3 &&
                          std::memcmp(c->h_vel.data(), vel, nc * 24) == 0;
    if (!same_vel) {
        c->h_vel.assign(vel, vel + nc * 3);
        if (int e = upload(c, c->d_vel, vel, nc * 24)) return e;
    }
    if (int e = upload(c, c->d_lw, lw, nc * 8)) return e;
    c->lw_gen++;
    c->umass_av = umass_av;
    c->aksmax = aksmax;
    c->have_medium = true;
    if (!same_vel)
        for (auto &s : c->incl) s.traj_ready = false;    // u depends on the velocity
    c->cells_line = -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));      // host arrays are only read during the call
    return 0;
}

extern "C" int rl_set_levels(rl_ctx *c, int nlev, const double *pops) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !c->nr || nlev < 2 || !pops) return fail(c, RL_E_ARG, "rl_set_levels: bad args");
    HIPCHK(c, hipSetDevice(c->device));
    c->nlev = nlev;
    c->cells_line = -1;
    if (int e = upload(c, c->d_pops, pops, (size_t)c->nr * c->nh * nlev * 8)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_set_lines(rl_ctx *c, int nl, const double *nu0, const double *A,
                            const double *gu, const double *gl, const int *lu, const int *ll) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || nl < 1 || !nu0 || !A || !gu || !gl || !lu || !ll)
        return fail(c, RL_E_ARG, "rl_set_lines: bad args");
    for (int i = 0; i < nl; i++) {                           // line.F:1739-1770
        if (lu[i] < 0 || ll[i] < 0 || (c->nlev && (lu[i] >= c->nlev || ll[i] >= c->nlev)))
            return fail(c, 13, "line %d: level index out of range", i);
        if (lu[i] <= ll[i]) return fail(c, 13, "line %d: upper level must follow lower", i);
        if (nu0[i] == 0.0) return fail(c, 13, "line %d: nu0 = 0", i);
    }
    c->nlines = nl;
    c->nu0.assign(nu0, nu0 + nl);
    c->A.assign(A, A + nl);
    c->gu.assign(gu, gu + nl);
    c->gl.assign(gl, gl + nl);
    c->lev_up.assign(lu, lu + nl);
    c->lev_low.assign(ll, ll + nl);
    c->have_dust = false;
    c->cells_line = -1;
    return 0;
}

extern "C" int rl_set_dust_line(rl_ctx *c, const double *jd, const double *ad) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !c->nr || !c->nlines || !jd || !ad)
        return fail(c, RL_E_ARG, "rl_set_dust_line: set grid and lines first");
    HIPCHK(c, hipSetDevice(c->device));
    size_t n = (size_t)c->nr * c->nh * c->nlines * 8;
    if (int e = upload(c, c->d_jd, jd, n)) return e;
    if (int e = upload(c, c->d_ad, ad, n)) return e;
    c->have_dust = true;
    c->cells_line = -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_set_star(rl_ctx *c, double rstar, int ncont, const double *nu,
                           const double *istar) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || rstar < 0.0 || ncont < 0 || (ncont > 0 && (!nu || !istar)))
        return fail(c, RL_E_ARG, "rl_set_star: bad args");
    if (c->nr && rstar >= c->rc[2]) return fail(c, 13, "star is bigger than inner grid radius");
    c->rstar = rstar;
    c->cont_nu.assign(nu, nu + ncont);
    c->istar.assign(istar, istar + ncont);
    for (auto &s : c->incl) s.traj_ready = s.counted = false;   // R* bounds the crossings
    c->x_owner.assign(c->x_owner.size(), -1);
    return 0;
}

extern "C" int rl_set_outer_bc(rl_ctx *c, int type) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c) return RL_E_ARG;
    if (type == 1) return fail(c, 13, "outer BC type 1 not allowed for telescope");
    if (type != 0 && type != 2 && type != 3)
        return fail(c, 13, "telescope: unknown outer boundary type %d", type);
    c->outer_bc = type;
    return 0;
}

extern "C" int rl_set_interstellar_field(rl_ctx *c, int ncont, const double *cont_nu,
                                         const double *isrf) {
    if (c) c->gen++;
    if (!c || ncont < 2 || !cont_nu || !isrf)
        return fail(c, RL_E_ARG, "rl_set_interstellar_field: bad args");
    c->isrf_nu.assign(cont_nu, cont_nu + ncont);
    c->isrf.assign(isrf, isrf + ncont);
    return 0;
}

// ---------------------------------------------------------------------
// rays: setup_rays_circular (telescope.F:715-1191) + ring edges
// (telescope.F:434-488) + flux weights (telescope.F:1386-1418)
// ---------------------------------------------------------------------
static int build_rayset(rl_ctx *c, RaySet &rs, double anginf, int nphi, int nrext, int dbdr) {
    const int nrref = 10;
    const double epsxyz = 1.0e2 * 1.0e-10, epsrrr = 1.0e3 * 1.0e-10;
    const int nr = c->nr;
    int nrrextra = nrext < 0 ? -nrext : nrext;
    auto R = [&](int i) { return c->rc[i + 1]; };
    if (nrrextra == 0) return fail(c, 13, "b_extra must be non-zero (main.F:628-635)");
    if (dbdr < 1 || nphi < 1) return fail(c, 13, "nphi and dbdr must be >= 1");
    if (std::fabs(anginf) < 1.0e-1) anginf = (double)0.1f * std::fabs(anginf) / anginf;
    if (c->rstar > R(1)) return fail(c, 91991, "star larger than the inner radius");
    const double theta0 = anginf + 1.0e-4;
    const double sinth0 = std::sin(theta0);
    const double dphi = 6.28318530718 / (1.0 * nphi);
    double phi = 0.5 * dphi;
    std::vector<double> xh(nphi), zh(nphi);
    for (int k = 0; k < nphi; k++) {
        zh[k] = -std::sin(phi) / sinth0;
        double zh02 = zh[k] * zh[k];
        double dum = 1.0 - zh02 * sinth0 * sinth0;
        dum = dum + epsxyz;
        if (dum < 0.0) return fail(c, 13, "ray direction setup failed");
        xh[k] = std::cos(phi) > 0.0 ? std::sqrt(dum) : -std::sqrt(dum);
        phi = phi + dphi;
    }
    const int nrings = nrrextra + nrref + (nr - 1) * dbdr;
    rs.theta0 = theta0;
    rs.costh0 = std::cos(theta0);                   // telescope.F:2884-2885
    rs.sinth0 = std::sin(theta0);
    rs.nrings = nrings;
    rs.nphi = nphi;
    rs.nrays = 1 + nrings * nphi;
    rs.x0.clear();
    rs.z0.clear();
    rs.ring_r.assign(nrings + 1, 0.0);
    rs.ring_ri.assign(nrings + 2, 0.0);
    // the rays of ring ir are r_ir (xh_k, zh_k), formed on the device
    int irad = 1;
    auto ring = [&](double r_c) { rs.ring_r[irad++] = r_c; };
    double ref1 = 0.5, ref2 = 0.0, rstar = c->rstar;
    for (int ix = 1; ix <= nrrextra + nrref; ix++) {
        if (ix <= nrrextra) {
            ring(((double)(ix - 1) * (R(1) - rstar) / nrrextra) + rstar);
        } else {
            ref2 = ref2 + ref1;
            ref1 = ref1 / 2;
            ring((((double)nrrextra + ref2) * (R(1) - rstar) / ((double)nrrextra + 1.0)) +
                 rstar);
        }
    }
    for (int ix = 1; ix <= nr - 1; ix++) {
        ring(R(ix) * (1.0 + epsrrr));
        if (dbdr > 1) {
            double dr = (R(ix + 1) - R(ix)) / (1.0 * dbdr);
            for (int iins = 1; iins <= dbdr - 1; iins++) ring(R(ix) + iins * dr);
        }
    }
    auto &rr = rs.ring_r;
    auto &ri = rs.ring_ri;
    ri[0] = 0.0;
    for (int k = 1; k <= nrings; k++) ri[k] = 0.5 * (rr[k] + rr[k - 1]);
    ri[nrings + 1] = R(nr);
    for (int k = 1; k <= nrings - 1; k++)
        if (rr[k] - ri[k] > 2 * (ri[k + 1] - rr[k])) ri[k] = rr[k] - 2 * (ri[k + 1] - rr[k]);
    if (ri[1] < rstar) {
        if (rr[1] < rstar) return fail(c, 1, "internal error in ray setup");
        ri[1] = rstar;
    }
    std::vector<double> w(nrings);
    for (int k = 1; k <= nrings; k++)
        w[k - 1] = 3.14159265359 * (ri[k + 1] * ri[k + 1] - ri[k] * ri[k]);
    rs.center_w = 3.14159265359 * (ri[1] * ri[1]);
    rs.raytab.assign(rs.ring_r.begin(), rs.ring_r.end());
    rs.raytab.insert(rs.raytab.end(), xh.begin(), xh.end());
    rs.raytab.insert(rs.raytab.end(), zh.begin(), zh.end());
    if (int e = upload(c, rs.d_raytab, rs.raytab.data(), rs.raytab.size() * 8)) return e;
    HIPCHK(c, rs.d_x0.ensure((size_t)rs.nrays * 8));
    HIPCHK(c, rs.d_z0.ensure((size_t)rs.nrays * 8));
    rl::launch_rays_circular(c->stream, nrings, nphi, rs.d_raytab.as<double>(),
                             rs.d_x0.as<double>(), rs.d_z0.as<double>());
    HIPCHK(c, hipGetLastError());
    rs.ring_w = w;
    if (int e = upload(c, rs.d_ringw, rs.ring_w.data(), w.size() * 8)) return e;
    rs.traj_ready = rs.counted = false;
    return 0;
}

static void keep_spares(rl_ctx *c);

extern "C" int rl_setup_rays_circular(rl_ctx *c, int nincl, const double *incl, int nphi,
                                      int nrext, int dbdr, int *n_rays, int *n_rings,
                                      double *ring_r, double *ring_edge) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !c->nr || nincl < 1 || !incl)
        return fail(c, RL_E_ARG, "rl_setup_rays_circular: set the grid first");
    HIPCHK(c, hipSetDevice(c->device));
    keep_spares(c);            // launches that still read them precede any reuse (one stream)
    c->incl.clear();
    c->x_owner.assign(c->x_owner.size(), -1);
    c->incl.resize(nincl);
    for (int i = 0; i < nincl; i++) {
        if (int e = build_rayset(c, c->incl[i], incl[i], nphi, nrext, dbdr)) return e;
        if (n_rays) n_rays[i] = c->incl[i].nrays;
        if (n_rings) n_rings[i] = c->incl[i].nrings;
    }
    if (ring_r) std::memcpy(ring_r, c->incl[0].ring_r.data(), c->incl[0].ring_r.size() * 8);
    if (ring_edge)
        std::memcpy(ring_edge, c->incl[0].ring_ri.data(), c->incl[0].ring_ri.size() * 8);
    return 0;
}

// setup_rays_rectang (telescope.F:2229-2468) for linespec_command 2: the
// pixel rays of an nx x ny image of the given size (cm), rotated by phioff
// and offset, ix outer and iy inner, then (addstar) the central ray for
// the added unresolved star
static int build_rayset_rect(rl_ctx *c, RaySet &rs, double anginf, int nx, int ny, double size_x,
                             double size_y, double phioff, double xoff, double yoff,
                             int addstar) {
    if (nx < 2 || ny < 2 || nx % 2 || ny % 2) return fail(c, 13, "nrx and nry must be even");
    const double szx = size_x * 0.5, szy = size_y * 0.5;             // telescope.F:209-212
    const double spx = szx / (nx * 0.5), spy = szy / (ny * 0.5);
    if (!(spx > 0.0) || !(spy > 0.0)) return fail(c, 13, "pixel size must be > 0");
    if (anginf < 1.0e-1) anginf = (double)0.1f;                      // :2294-2303
    const double theta0 = anginf + 1.0e-4;
    const double sinth0 = std::sin(theta0);
    rs = RaySet();
    rs.rect = true;
    rs.theta0 = theta0;
    rs.costh0 = std::cos(theta0);
    rs.sinth0 = sinth0;
    rs.nx = nx; rs.ny = ny; rs.spx = spx; rs.spy = spy;
    const int n = nx * ny + (addstar > 0 ? 1 : 0);
    rs.nrays = n;
    rs.x0.assign(n, 0.0);
    rs.z0.assign(n, 0.0);
    rs.bpix.assign(n, 0.0);
    const int nxh = nx / 2, nyh = ny / 2;
    int ir = 0;
    for (int ix = 1; ix <= nx; ix++) {
        for (int iy = 1; iy <= ny; iy++) {
            double x_c = (ix - nxh - 0.5) * spx - xoff;
            const double y_c = (iy - nyh - 0.5) * spy - yoff;
            const double r_c = std::sqrt(x_c * x_c + y_c * y_c);
            if (x_c == 0.0) x_c = x_c + 0.001 * spx;
            double p_c = std::atan(y_c / x_c) - phioff;
            if (x_c < 0.0) p_c = p_c + 3.14159265359;
            while (p_c < 0.0) p_c = p_c + 6.28318530718;
            while (p_c >= 6.28318530718) p_c = p_c - 6.28318530718;
            const double zh0 = -std::sin(p_c) / sinth0;
            const double zh02 = zh0 * zh0;
            double dum = 1.0 - zh02 * sinth0 * sinth0;
            dum = dum + 1e-4;
            if (dum < 0.0) return fail(c, 13, "setup_rays_rectang: ray direction");
            const double xh0 = std::cos(p_c) > 0.0 ? std::sqrt(dum) : -std::sqrt(dum);
            rs.x0[ir] = r_c * xh0;
            rs.z0[ir] = r_c * zh0;
            rs.bpix[ir] = r_c;
            ir++;
        }
    }
    if (addstar > 0) {                                               // :2349-2400
        if (xoff != 0.0 || yoff != 0.0)
            return fail(c, 13, "the added star needs a star-centred image (offsets 0)");
        rs.central = ir;
        if (spx * spx + spy * spy > c->rstar * c->rstar) rs.starunres = 1;
    }
    // rays the imager traces whose impact parameter is within the star
    const double rmax = c->rc[c->nr + 1], costh02 = rs.costh0 * rs.costh0;
    for (int i = 0; i < n; i++) {
        const double bimpact = std::sqrt(rs.x0[i] * rs.x0[i] + rs.z0[i] * rs.z0[i] * (1.0 - costh02));
        if (rs.bpix[i] < 0.999 * rmax && bimpact <= c->rstar) rs.star_rays.push_back(i);
    }
    if (int e = upload(c, rs.d_x0, rs.x0.data(), (size_t)n * 8)) return e;
    if (int e = upload(c, rs.d_z0, rs.z0.data(), (size_t)n * 8)) return e;
    if (int e = upload(c, rs.d_bpix, rs.bpix.data(), (size_t)n * 8)) return e;
    if (!rs.star_rays.empty())
        if (int e = upload(c, rs.d_star_rays, rs.star_rays.data(), rs.star_rays.size() * 4))
            return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_setup_rays_rectang(rl_ctx *c, int nincl, const double *incl, int nx, int ny,
                                     double size_x, double size_y, double phioff, double xoff,
                                     double yoff, int addstar, int *n_rays) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !c->nr || nincl < 1 || !incl)
        return fail(c, RL_E_ARG, "rl_setup_rays_rectang: set the grid first");
    HIPCHK(c, hipSetDevice(c->device));
    keep_spares(c);            // launches that still read them precede any reuse (one stream)
    c->incl.clear();
    c->x_owner.assign(c->x_owner.size(), -1);
    c->incl.resize(nincl);
    for (int i = 0; i < nincl; i++) {
        if (int e = build_rayset_rect(c, c->incl[i], incl[i], nx, ny, size_x, size_y, phioff,
                                      xoff, yoff, addstar))
            return e;
        if (n_rays) n_rays[i] = c->incl[i].nrays;
    }
    return 0;
}

// ---------------------------------------------------------------------
// trace
// ---------------------------------------------------------------------
// upper bound of the nodes of one ray: every R crossing, every cone
// crossing and every extra point (telescope.F:3279, 3196, RAYEXPT),
// rounded up to 8 records so that each ray's crossing slot starts on a
// 128-byte line (the crossing pass writes whole lines)
static int traj_nmax(const rl_ctx *c) {
    return (2 * (c->nr + 1) + 2 * c->nh + 40 + 7) / 8 * 8;
}

static int check_err(rl_ctx *c) {
    int e = 0;
    HIPCHK(c, hipMemcpyAsync(&e, c->d_err.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (e) return fail(c, e, "device stop code %d", e);
    return 0;
}

// device bytes a ray set holds: nodes and, once built, the adjoint's
// stencil transpose
static size_t traj_bytes(const rl_ctx *c, const RaySet &rs) {
    size_t b = (size_t)rs.nnodes * sizeof(GeoNode) + (size_t)(rs.nrays + 1) * 8;
    if (c->precision == RL_FP32 && !rs.rect)
        b += (size_t)rs.nnodes * (sizeof(rl::SegHot) + sizeof(rl::SegAux));
    if (rs.st_ready) b += (size_t)rs.nnodes * 4 * 12 + ((size_t)c->nr * c->nh + 1) * 4;
    return b;
}

// device bytes a ray set holds now
static size_t held_bytes(const RaySet &rs) {
    return rs.d_nodes.n + rs.d_st_vals.n + rs.d_st_w.n + rs.d_st_off.n + rs.d_hot.n + rs.d_aux.n;
}

static size_t resident_traj(const rl_ctx *c) {
    size_t b = 0;
    for (const RaySet &r : c->incl) b += held_bytes(r);
    return b;
}

// Spare node buffers: those of ray sets that a setup call replaced or that
// were evicted to make room, kept for the next builds of the same sizes
// (allocating and freeing tens of GB costs seconds).  They live only
// between an eviction (or setup) and the builds that follow it: a build
// that finds none of a fitting size frees them all before it allocates,
// and release_cross_scratch frees what no build took.
static void keep_spares(rl_ctx *c) {
    for (RaySet &r : c->incl)
        for (DevBuf *d : {&r.d_nodes, &r.d_hot, &r.d_aux})
            if (d->p) c->spare.push_back(std::move(*d));
}

static void free_spares(rl_ctx *c) {
    if (c->spare.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    c->spare.clear();
}

// b holds at least `bytes`: a spare buffer if one fits (the smallest of at
// most twice the size), else a new allocation (the spares freed first)
static hipError_t ensure_spare(rl_ctx *c, DevBuf &b, size_t bytes) {
    if (b.p && b.n >= bytes) return hipSuccess;
    int best = -1;
    for (int i = 0; i < (int)c->spare.size(); i++) {
        const size_t n = c->spare[i].n;
        if (n >= bytes && n <= 2 * bytes + 4096 && (best < 0 || n < c->spare[best].n)) best = i;
    }
    if (best < 0) {
        free_spares(c);
        return b.ensure(bytes);
    }
    b = std::move(c->spare[best]);
    c->spare.erase(c->spare.begin() + best);
    return hipSuccess;
}

// the memory budget: set by rl_set_memory_budget, else 60% of the device
// memory free at the context's first trace plus what the context held
// then (fixed, so that repeated calls plan alike)
static size_t ctx_budget(rl_ctx *c) {
    if (c->mem_budget) return c->mem_budget;
    if (c->auto_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
        c->auto_budget = (size_t)(0.6 * (double)fr) + resident_traj(c) + c->d_rec.n +
                         c->d_image.n + c->d_partial.n + c->d_ckpt.n + c->d_arec.n;
    }
    return c->auto_budget;
}

// The forward's batch buffers and the adjoint's chunk buffers are kept
// between calls for reuse; one kind is released when the other kind of
// call starts and it holds more than a tenth of the budget.
static void release_forward_work(rl_ctx *c) {
    if ((c->d_rec.n + c->d_cmask.n) * 10 > ctx_budget(c)) {
        (void)hipStreamSynchronize(c->stream);
        c->d_rec = DevBuf();
        c->d_cmask = DevBuf();
    }
}

static void release_adjoint_work(rl_ctx *c) {
    if ((c->d_arec.n + c->d_ckpt.n + c->d_gnode.n + c->d_gcell.n) * 10 > ctx_budget(c)) {
        (void)hipStreamSynchronize(c->stream);
        c->d_arec = DevBuf();
        c->d_ckpt = DevBuf();
        c->d_gnode = DevBuf();
        c->d_gcell = DevBuf();
    }
}

static void evict_traj(rl_ctx *c, RaySet &rs) {
    for (DevBuf *d : {&rs.d_nodes, &rs.d_hot, &rs.d_aux})
        if (d->p) c->spare.push_back(std::move(*d));
    rs.seg_ready = false;
    rs.d_st_vals = DevBuf();
    rs.d_st_w = DevBuf();
    rs.d_st_off = DevBuf();
    rs.traj_ready = false;
    rs.st_ready = false;
}

// evict the least recently used ray sets other than those in keep until
// the resident trajectories plus need fit the budget (or none is left)
static int make_room(rl_ctx *c, size_t need, const std::vector<int> &keep) {
    const size_t budget = ctx_budget(c);
    for (;;) {
        if (resident_traj(c) + need <= budget) return 0;
        int victim = -1;
        for (int i = 0; i < (int)c->incl.size(); i++) {
            const RaySet &r = c->incl[i];
            if (!held_bytes(r) || std::find(keep.begin(), keep.end(), i) != keep.end()) continue;
            if (victim < 0 || r.used < c->incl[victim].used) victim = i;
        }
        if (victim < 0) return 0;
        HIPCHK(c, hipStreamSynchronize(c->stream));   // no launch may still read it
        evict_traj(c, c->incl[victim]);
    }
}

// ray crossings of a ray set (k_cross) into the context's scratch, and
// the node counts: off / nnodes
// ray crossings (k_cross) of up to CROSS_SETS_MAX ray sets in one launch,
// each into its own scratch slot, and their node counts: off / nnodes
static int cross_sets(rl_ctx *c, const std::vector<int> &idxs) {
    const int ns = (int)idxs.size();
    if (ns < 1 || ns > rl::CROSS_SETS_MAX) return fail(c, RL_E_ARG, "crossing: %d ray sets", ns);
    rl::GridDev g{c->nr, 2 * c->nh, c->d_rc.as<double>(), c->d_tc.as<double>(), c->rstar,
                  c->t_sorted};
    const int iyeq = c->nh, nmax = traj_nmax(c);
    if ((int)c->x_scr.size() < ns) {
        c->x_scr.resize(ns);
        c->x_cross.resize(ns);
        c->x_owner.resize(ns, -1);
    }
    rl::CrossSets sets{};
    sets.n = ns;
    if ((int)c->x_cnt.size() < ns) c->x_cnt.resize(ns);
    std::vector<DevBuf> &cnt = c->x_cnt;
    for (int k = 0; k < ns; k++) {
        RaySet &rs = c->incl[idxs[k]];
        const int n = rs.nrays;
        c->x_owner[k] = -1;
        HIPCHK(c, c->x_scr[k].ensure((size_t)(iyeq + 1) * n * 20));
        HIPCHK(c, c->x_cross[k].ensure((size_t)n * nmax * sizeof(rl::CrossRec)));
        HIPCHK(c, cnt[k].ensure((size_t)n * 4));
        rl::CrossSet &q = sets.s[k];
        q.nrays = n;
        q.central_blend = !rs.rect && rs.ring_ri[1] > 0.0 ? 1 : 0;
        q.bpix = rs.rect ? rs.d_bpix.as<double>() : nullptr;
        q.bmax = 0.999 * c->rc[c->nr + 1];               // telescope.F:2116-2118
        q.x0 = rs.d_x0.as<double>();
        q.z0 = rs.d_z0.as<double>();
        q.theta0 = rs.theta0; q.costh0 = rs.costh0; q.sinth0 = rs.sinth0;
        q.s1 = c->x_scr[k].as<double>();
        q.s2 = q.s1 + (size_t)(iyeq + 1) * n;
        q.iy = reinterpret_cast<int *>(q.s2 + (size_t)(iyeq + 1) * n);
        q.count = cnt[k].as<int>();
        q.cr = c->x_cross[k].as<rl::CrossRec>();
    }
    HIPCHK(c, hipMemsetAsync(c->d_err.p, 0, 4, c->stream));
    rl::launch_cross(c->stream, g, c->d_tan2.as<double>(), sets, nmax, c->d_err.as<int>());
    HIPCHK(c, hipGetLastError());
    std::vector<std::vector<int>> h(ns);
    for (int k = 0; k < ns; k++) {
        h[k].resize(sets.s[k].nrays);
        HIPCHK(c, hipMemcpyAsync(h[k].data(), cnt[k].p, (size_t)sets.s[k].nrays * 4,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    if (int e = check_err(c)) return e;
    for (int k = 0; k < ns; k++) {
        RaySet &rs = c->incl[idxs[k]];
        const int n = rs.nrays;
        rs.off.assign(n + 1, 0);
        for (int i = 0; i < n; i++) rs.off[i + 1] = rs.off[i] + h[k][i];
        rs.nnodes = rs.off[n];
        if (int e = upload(c, rs.d_off, rs.off.data(), (size_t)(n + 1) * 8)) return e;
        rs.counted = true;
        c->x_owner[k] = idxs[k];
    }
    return 0;
}

static int cross_traj(rl_ctx *c, int idx) { return cross_sets(c, std::vector<int>{idx}); }

// node counts of a ray set (without building its nodes)
static int count_traj(rl_ctx *c, int idx) {
    if (c->incl[idx].counted) return 0;
    if (int e = cross_traj(c, idx)) return e;
    return 0;
}

// After the builds of a call: the spares no build took are freed; the
// scratch of k_cross (32 GB on the 1024x512 disk) is kept for the next
// build unless it and the resident trajectories exceed the budget.
static void release_cross_scratch(rl_ctx *c) {
    free_spares(c);
    size_t b = 0;
    for (size_t k = 0; k < c->x_scr.size(); k++) b += c->x_scr[k].n + c->x_cross[k].n;
    if (b + resident_traj(c) > ctx_budget(c)) {
        (void)hipStreamSynchronize(c->stream);
        c->x_scr.clear();
        c->x_cross.clear();
        c->x_cnt.clear();
        c->x_owner.clear();
    }
}

// the nodes of a ray set, built if not resident; other ray sets (not in
// keep) are evicted first when the budget requires
static int build_traj_keep(rl_ctx *c, int idx, const std::vector<int> &keep) {
    RaySet &rs = c->incl[idx];
    rs.used = ++c->tick;
    if (rs.traj_ready) return 0;
    if (int e = count_traj(c, idx)) return e;
    std::vector<int> k2 = keep;
    k2.push_back(idx);
    const size_t want = traj_bytes(c, rs), have = held_bytes(rs);
    if (int e = make_room(c, want > have ? want - have : 0, k2)) return e;
    int slot = (int)(std::find(c->x_owner.begin(), c->x_owner.end(), idx) - c->x_owner.begin());
    if (slot == (int)c->x_owner.size()) {
        if (int e = cross_traj(c, idx)) return e;
        slot = 0;
    }
    rl::GridDev g{c->nr, 2 * c->nh, c->d_rc.as<double>(), c->d_tc.as<double>(), c->rstar,
                  c->t_sorted};
    const int nmax = traj_nmax(c);
    const rl::CrossRec *cr = c->x_cross[slot].as<rl::CrossRec>();
    HIPCHK(c, ensure_spare(c, rs.d_nodes, (size_t)rs.nnodes * sizeof(GeoNode)));
    // FP32 circular sets: the segment records in the same pass
    const bool seg = c->precision == RL_FP32 && !rs.rect && c->d_lw.n;
    if (seg) {
        HIPCHK(c, ensure_spare(c, rs.d_hot, (size_t)rs.nnodes * sizeof(rl::SegHot)));
        HIPCHK(c, ensure_spare(c, rs.d_aux, (size_t)rs.nnodes * sizeof(rl::SegAux)));
        HIPCHK(c, c->d_vlw.ensure((size_t)c->nr * c->nh * sizeof(rl::VelLw)));
        HIPCHK(c, c->d_xz.ensure((size_t)rs.nrays * sizeof(rl::RayXZ)));
    }
    rl::launch_geo(c->stream, g, c->d_vel.as<double>(), c->nh, rs.nrays, nmax,
                   rs.d_off.as<long long>(), rs.d_x0.as<double>(), rs.d_z0.as<double>(),
                   rs.costh0, rs.sinth0, cr, rs.d_nodes.as<GeoNode>(),
                   c->d_err.as<int>(), rs.nnodes, seg ? c->d_lw.as<double>() : nullptr,
                   seg ? rs.d_hot.as<rl::SegHot>() : nullptr,
                   seg ? rs.d_aux.as<rl::SegAux>() : nullptr,
                   seg ? c->d_vlw.as<rl::VelLw>() : nullptr,
                   seg ? c->d_xz.as<rl::RayXZ>() : nullptr);
    HIPCHK(c, hipGetLastError());
    if (rs.rect && !rs.star_rays.empty())
        rl::launch_star_flags(c->stream, (int)rs.star_rays.size(), rs.d_star_rays.as<int>(),
                              rs.d_off.as<long long>(), rs.d_nodes.as<GeoNode>(), c->nh);
    HIPCHK(c, hipGetLastError());
    if (int e = check_err(c)) return e;
    rs.traj_ready = true;
    rs.st_ready = false;
    rs.seg_ready = seg;
    rs.seg_lw_gen = c->lw_gen;
    return 0;
}

// the FP32 segment records of a built ray set, current with the line widths
static int ensure_segrec(rl_ctx *c, RaySet &rs) {
    if (c->precision != RL_FP32 || rs.rect) return 0;
    if (!rs.traj_ready) return fail(c, 13, "segment records: trajectories not built");
    if (rs.seg_ready && rs.seg_lw_gen == c->lw_gen) return 0;
    HIPCHK(c, rs.d_hot.ensure((size_t)rs.nnodes * sizeof(rl::SegHot)));
    HIPCHK(c, rs.d_aux.ensure((size_t)rs.nnodes * sizeof(rl::SegAux)));
    rl::launch_segrec(c->stream, rs.nnodes, rs.d_off.as<long long>(), rs.nrays,
                      rs.d_nodes.as<GeoNode>(), c->d_lw.as<double>(),
                      rs.d_hot.as<rl::SegHot>(), rs.d_aux.as<rl::SegAux>());
    HIPCHK(c, hipGetLastError());
    rs.seg_ready = true;
    rs.seg_lw_gen = c->lw_gen;
    return 0;
}

static int build_traj(rl_ctx *c, RaySet &rs) {
    return build_traj_keep(c, (int)(&rs - c->incl.data()), std::vector<int>());
}

// the stencil transpose of a ray set's nodes (once per trajectories)
// the stencil transpose of nodes [n0, n1) of a ray set (cell -> the
// entries 4k+j, k relative to n0), into vals / w / off
static int build_transpose_range(rl_ctx *c, RaySet &rs, long long n0, long long n1,
                                 DevBuf &vals, DevBuf &w, DevBuf &off) {
    const int ncell = c->nr * c->nh;
    const long long nn = n1 - n0;
    const size_t nent = 4 * (size_t)nn;
    if (nent >= (size_t)1 << 31)
        return fail(c, 13, "adjoint: %lld nodes exceed the stencil index range", nn);
    DevBuf keys_a, keys_b, vals_a, temp;
    HIPCHK(c, keys_a.ensure(nent * 4));
    HIPCHK(c, keys_b.ensure(nent * 4));
    HIPCHK(c, vals_a.ensure(nent * 4));
    HIPCHK(c, vals.ensure(nent * 4));
    HIPCHK(c, w.ensure(nent * 8));
    HIPCHK(c, off.ensure(((size_t)ncell + 1) * 4));
    const GeoNode *geo = rs.d_nodes.as<GeoNode>() + n0;
    size_t tb = 0;
    if (rl::build_stencil_transpose(c->stream, nn, geo, ncell, keys_a.as<int>(),
                                    keys_b.as<int>(), vals_a.as<int>(), vals.as<int>(),
                                    w.as<double>(), off.as<int>(), nullptr, &tb))
        return fail(c, RL_E_HIP, "stencil transpose: sort size query failed");
    HIPCHK(c, temp.ensure(tb));
    if (int e = rl::build_stencil_transpose(c->stream, nn, geo, ncell, keys_a.as<int>(),
                                            keys_b.as<int>(), vals_a.as<int>(), vals.as<int>(),
                                            w.as<double>(), off.as<int>(), temp.p, &tb))
        return fail(c, RL_E_HIP, "stencil transpose: sort failed (%d)", e);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));      // temporaries are freed on return
    return 0;
}

// the whole ray set's transpose, kept with its nodes
static int build_transpose(rl_ctx *c, RaySet &rs) {
    if (rs.st_ready) return 0;
    if (int e = build_transpose_range(c, rs, 0, rs.nnodes, rs.d_st_vals, rs.d_st_w, rs.d_st_off))
        return e;
    rs.st_ready = true;
    return 0;
}

static double star_at(const rl_ctx *c, double freq) {      // line.F:3821-3843
    const int n = (int)c->cont_nu.size();
    if (n < 2) return 0.0;
    const double *x = c->cont_nu.data() - 1, *s = c->istar.data() - 1;
    bool asc = x[n] > x[1];
    int lo = 0, hi = n + 1;
    while (hi - lo != 1) {
        int m = (hi + lo) >> 1;
        if ((freq > x[m]) == asc) lo = m; else hi = m;
    }
    if (lo == 0 || lo == n) return 0.0;
    double wgt = (freq - x[lo]) / (x[lo + 1] - x[lo]);
    return (1.0 - wgt) * s[lo] + wgt * s[lo + 1];
}

// Outer boundary intensity of every channel (charintline,
// telescope.F:3996-4026): 0 (type 0), the CMB at the line centre (type 2),
// or the interstellar field at each channel frequency nu0 + dnu
// (line_init_passband_outerbc, line.F:3878-3900) on the circular-image
// path; the pos-vel path (calc_write_line_posvel) never fills the field,
// so it starts from 0 there.  The reference tests the upper end of the
// continuum interval against freq_nr, which line_setup_passband has just
// set to the channel count (line.F:505), so a channel in interval nfr
// gets 0; an interval that ends at the last continuum point reads the
// (zero) point after it.
static void outer_bc_channels(const rl_ctx *c, int line, int nfr, const double *dnu, bool rect,
                              double *out) {
    const double nu0s = c->nu0[line], nu0 = std::fabs(nu0s);
    double v = 0.0;
    if (c->outer_bc == 2)
        v = 1.47455253991e-47 * (nu0s * nu0s * nu0s) / (std::exp(4.7991598e-11 * nu0s / 2.728) - 1.0);
    for (int k = 0; k < nfr; k++) out[k] = v;
    if (c->outer_bc != 3 || rect) return;
    const int n = (int)c->isrf_nu.size();
    const double *x = c->isrf_nu.data() - 1, *f = c->isrf.data() - 1;
    const bool asc = x[n] > x[1];
    for (int k = 0; k < nfr; k++) {
        const double freq = nu0 + dnu[k];
        int lo = 0, hi = n + 1;
        while (hi - lo != 1) {
            const int m = (hi + lo) >> 1;
            if ((freq > x[m]) == asc) lo = m; else hi = m;
        }
        if (lo == 0 || lo == nfr) continue;
        const double x1 = lo < n ? x[lo + 1] : 0.0, f1 = lo < n ? f[lo + 1] : 0.0;
        const double wgt = (freq - x[lo]) / (x1 - x[lo]);
        out[k] = (1.0 - wgt) * f[lo] + wgt * f1;
    }
}

// Per-item set-up shared by the forward and the adjoint: trajectories of
// the item's inclination, the line's cell table, the channel tables
// (line_setup_passband, line.F:462-469; star per channel, line.F:3821-3843)
// and the line constants (line.F:1778-1787, 2301, 4578-4588).
struct ItemSetup {
    RaySet *rs = nullptr;
    rl::LineConst lc;
};

static int setup_item(rl_ctx *c, const rl_item &it, int ii, ItemSetup &out, hipEvent_t ev0,
                      hipEvent_t ev1) {
    if (it.line < 0 || it.line >= c->nlines || it.incl < 0 ||
        it.incl >= (int)c->incl.size() || it.nfr < 2 || !(it.vmax_kms > 0.0) ||
        (it.skip_mode != RL_SKIP_REFERENCE && it.skip_mode != RL_SKIP_NONE))
        return fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", ii);
    RaySet &rs = c->incl[it.incl];
    out.rs = &rs;
    if (ev0) HIPCHK(c, hipEventRecord(ev0, c->stream));
    if (int e = build_traj(c, rs)) return e;
    if (ev1) HIPCHK(c, hipEventRecord(ev1, c->stream));
    // cell table for this line
    const double molpg = 1.0 / (c->umass_av * 1.6726e-24);
    const int ncell = c->nr * c->nh;
    HIPCHK(c, c->d_cells.ensure((size_t)ncell * sizeof(CellRec)));
    if (c->cells_line != it.line) {
        rl::launch_cells(c->stream, ncell, c->nlev, c->nlines, it.line,
                         c->lev_up[it.line], c->lev_low[it.line], molpg,
                         c->d_rho.as<double>(), c->d_abund.as<double>(),
                         c->d_lw.as<double>(), c->d_pops.as<double>(),
                         c->d_jd.as<double>(), c->d_ad.as<double>(),
                         c->d_cells.as<CellRec>());
        c->cells_line = it.line;
    }
    // channel tables: line_setup_passband (line.F:462-469)
    const int nfr = it.nfr;
    const double nu0s = c->nu0[it.line], nu0 = std::fabs(nu0s);
    const double passb = 3.33567e-6 * nu0 * it.vmax_kms;
    const double nu1 = 0.0 - passb, dnu = 2.0 * passb / (nfr - 1.0);
    // kept in the context: the copy below is asynchronous
    std::vector<double> &chan = c->h_chan;
    chan.assign(4 * (size_t)nfr, 0.0);
    for (int k = 0; k < nfr; k++) {
        double dn = nu1 + k * dnu;
        chan[k] = dn;
        chan[nfr + k] = dn / nu0s;
        // the pos-vel path of the reference leaves the star's passband
        // spectrum unset (calc_write_line_posvel calls no
        // line_init_passband_starbc, cf. telescope.F:1557): 0 there
        chan[2 * nfr + k] = rs.rect ? 0.0 : star_at(c, nu0 + dn);
    }
    outer_bc_channels(c, it.line, nfr, chan.data(), rs.rect, chan.data() + 3 * nfr);
    if (int e = upload(c, c->d_chan, chan.data(), chan.size() * 8)) return e;
    rl::LineConst &lc = out.lc;
    lc.nu0 = nu0s;
    lc.c_aa = 3.33567e-6 * nu0;
    lc.hknu = 5.27296241956e-28 * nu0s;
    lc.A = c->A[it.line];
    lc.Bud = 6.78171833781e46 * lc.A / (nu0s * nu0s * nu0s);      // line.F:1778
    lc.Bdu = lc.Bud * (c->gu[it.line] / c->gl[it.line]);
    lc.starfract = 0.0;
    if (!rs.rect) {
        double q = c->rstar / rs.ring_ri[1];
        lc.starfract = q * q;
        if (rs.ring_ri[1] < c->rstar) return fail(c, 124, "central beam resolves the star");
    }
    return 0;
}

// line constants and channel tables of one item (line_setup_passband,
// line.F:462-469; star per channel, line.F:3821-3843; prepare_lines,
// line.F:1778-1787); chan receives dnu, dnu/nu0, I_star, I_outer [4*nfr]
static int item_consts(rl_ctx *c, const rl_item &it, const RaySet &rs, rl::LineConst &lc,
                       double *chan) {
    const int nfr = it.nfr;
    const double nu0s = c->nu0[it.line], nu0 = std::fabs(nu0s);
    const double passb = 3.33567e-6 * nu0 * it.vmax_kms;
    const double nu1 = 0.0 - passb, dnu = 2.0 * passb / (nfr - 1.0);
    for (int k = 0; k < nfr; k++) {
        double dn = nu1 + k * dnu;
        chan[k] = dn;
        chan[nfr + k] = dn / nu0s;
        chan[2 * nfr + k] = star_at(c, nu0 + dn);
    }
    outer_bc_channels(c, it.line, nfr, chan, rs.rect, chan + 3 * nfr);
    lc.nu0 = nu0s;
    lc.c_aa = 3.33567e-6 * nu0;
    lc.hknu = 5.27296241956e-28 * nu0s;
    lc.A = c->A[it.line];
    lc.Bud = 6.78171833781e46 * lc.A / (nu0s * nu0s * nu0s);
    lc.Bdu = lc.Bud * (c->gu[it.line] / c->gl[it.line]);
    double q = c->rstar / rs.ring_ri[1];
    lc.starfract = q * q;
    if (rs.ring_ri[1] < c->rstar) return fail(c, 124, "central beam resolves the star");
    return 0;
}

// Device bytes one item needs in a batch (node records, image, ring
// partial sums, channel tables, flux, ray info).
static size_t item_bytes(const rl_ctx *c, const rl_item &it, const RaySet &rs) {
    const size_t rsz = c->precision == RL_FP64 ? sizeof(rl::NodeRec<double>) : sizeof(rl::NodeLine);
    return (size_t)rs.nnodes * rsz + (size_t)rs.nrays * it.nfr * 8 +
           (size_t)rs.nrings * it.nfr * 8 + (size_t)it.nfr * 32 + (size_t)rs.nrays * 16 + 64;
}

static long long tiles_of(int cw, int nrays, int nfr) {
    if (cw == 0) return ((long long)nrays * nfr + 63) / 64;
    const int rows = 64 / cw;
    return (long long)((nrays + rows - 1) / rows) * ((nfr + cw - 1) / cw);
}

// batches: consecutive items whose buffers, with the nodes of their
// inclinations, fit the budget (at least one item per batch; node counts
// known).  Returns the largest batch's bytes.
static size_t cut_batches(rl_ctx *c, int nitems, const rl_item *items,
                          std::vector<std::pair<int, int>> &cuts) {
    const size_t budget = ctx_budget(c);
    size_t most = 0;
    int i0 = 0;
    while (i0 < nitems) {
        size_t used = 0;
        std::vector<int> incs;
        int i1 = i0;
        while (i1 < nitems) {
            const int in = items[i1].incl;
            size_t b = item_bytes(c, items[i1], c->incl[in]);
            if (std::find(incs.begin(), incs.end(), in) == incs.end())
                b += traj_bytes(c, c->incl[in]);
            if (i1 > i0 && (used + b > budget || i1 - i0 >= 65535)) break;   // grid y/z limit
            used += b;
            if (std::find(incs.begin(), incs.end(), in) == incs.end()) incs.push_back(in);
            i1++;
        }
        cuts.emplace_back(i0, i1);
        most = std::max(most, used);
        i0 = i1;
    }
    return most;
}

// One batch: items [i0, i1) of the call, all resident at once.
static int run_batch(rl_ctx *c, const rl_item *items, int i0, int i1, double *flux,
                     double *image, unsigned char *cmask, unsigned *warn, size_t fo,
                     size_t io) {
    const int nb = i1 - i0;
    const int fp64 = c->precision == RL_FP64;
    const size_t rsz = fp64 ? sizeof(rl::NodeRec<double>) : sizeof(rl::NodeLine);
    const int ncell = c->nr * c->nh;
    const double molpg = 1.0 / (c->umass_av * 1.6726e-24);
    // host layout of the batch
    std::vector<rl::ItemDev> hi(nb);
    std::vector<int> slot_line;
    std::vector<double> &chan = c->h_chan;
    size_t nchan = 0;
    for (int k = 0; k < nb; k++) nchan += 4 * (size_t)items[i0 + k].nfr;
    chan.assign(nchan, 0.0);
    long long rec0 = 0, img0 = 0, chan0 = 0, part0 = 0, flux0 = 0;
    int rinfo0 = 0, max_rays = 0, max_nfr = 0, max_rings = 0, any_skip = 0;
    for (int k = 0; k < nb; k++) {
        const rl_item &it = items[i0 + k];
        RaySet &rs = c->incl[it.incl];
        rl::ItemDev &d = hi[k];
        if (int e = item_consts(c, it, rs, d.lc, chan.data() + chan0)) return e;
        int slot = -1;
        for (size_t q = 0; q < slot_line.size(); q++)
            if (slot_line[q] == it.line) slot = (int)q;
        if (slot < 0) { slot = (int)slot_line.size(); slot_line.push_back(it.line); }
        d.nrays = rs.nrays; d.nfr = it.nfr; d.nrings = rs.nrings; d.nphi = rs.nphi;
        d.skip_mode = it.skip_mode; d.cslot = slot;
        if (int e = ensure_segrec(c, rs)) return e;
        d.off = rs.d_off.as<long long>();
        d.geo = rs.d_nodes.as<GeoNode>();
        d.hot_off = (long long)(uintptr_t)rs.d_hot.p;     // made relative below
        d.aux_off = (long long)(uintptr_t)rs.d_aux.p;
        d.rec0 = rec0; d.img0 = img0; d.chan0 = chan0; d.part0 = part0; d.flux0 = flux0;
        d.rinfo0 = rinfo0;
        d.aks2 = 2.0 * (c->aksmax / 2.99792458e5);
        d.ring_w = rs.d_ringw.as<double>();
        d.center_w = rs.center_w;
        rec0 += rs.nnodes; img0 += (long long)rs.nrays * it.nfr; chan0 += 4 * it.nfr;
        part0 += (long long)rs.nrings * it.nfr; flux0 += it.nfr; rinfo0 += rs.nrays;
        max_rays = std::max(max_rays, rs.nrays);
        max_nfr = std::max(max_nfr, it.nfr);
        max_rings = std::max(max_rings, rs.nrings);
        any_skip |= it.skip_mode == RL_SKIP_REFERENCE;
        c->counts[0] += rs.nrays;
        c->counts[1] += rs.nnodes;
        c->counts[3] += (long long)rs.nrays * it.nfr;
    }
    // device buffers
    HIPCHK(c, c->d_rec.ensure((size_t)rec0 * rsz));
    HIPCHK(c, c->d_image.ensure((size_t)img0 * 8));
    HIPCHK(c, c->d_partial.ensure((size_t)part0 * 8));
    HIPCHK(c, c->d_flux.ensure((size_t)flux0 * 8));
    HIPCHK(c, c->d_rinfo.ensure((size_t)rinfo0 * sizeof(RayInfo)));
    HIPCHK(c, c->d_cells.ensure((size_t)slot_line.size() * ncell * sizeof(CellRec)));
    HIPCHK(c, c->d_ovs.ensure((size_t)nb * 8));
    HIPCHK(c, c->d_maser.ensure((size_t)((nb + 1) / 2) * 8));
    HIPCHK(c, c->d_counter.ensure(8));
    HIPCHK(c, c->d_items.ensure((size_t)nb * sizeof(rl::ItemDev)));
    HIPCHK(c, c->d_tile0.ensure((size_t)(nb + 1) * 8));
    if (cmask) HIPCHK(c, c->d_cmask.ensure((size_t)img0));
    if (int e = upload(c, c->d_chan, chan.data(), nchan * 8)) return e;
    c->cells_line = -1;       // d_cells is rewritten below for this batch's lines
    for (auto &d : hi) {      // segment records relative to the record buffer
        d.hot_off -= (long long)(uintptr_t)c->d_rec.p;
        d.aux_off -= (long long)(uintptr_t)c->d_rec.p;
    }
    c->h_items = hi;
    if (int e = upload(c, c->d_items, c->h_items.data(), (size_t)nb * sizeof(rl::ItemDev))) return e;
    rl::BatchLaunch L;
    L.nitems = nb; L.max_rays = max_rays; L.max_nfr = max_nfr; L.max_rings = max_rings;
    L.ncell = ncell; L.any_skip = any_skip;
    L.items = c->d_items.as<rl::ItemDev>();
    L.cells = c->d_cells.as<CellRec>();
    L.rec = c->d_rec.p;
    L.rinfo = c->d_rinfo.as<RayInfo>();
    L.chan = c->d_chan.as<double>();
    L.image = c->d_image.as<double>();
    L.partial = c->d_partial.as<double>();
    L.flux = c->d_flux.as<double>();
    L.d2 = (1.0 * 3.08572e18) * (1.0 * 3.08572e18);
    L.tile0 = c->d_tile0.as<long long>();
    L.ovs = c->d_ovs.as<unsigned long long>();
    L.counter = c->d_counter.as<unsigned long long>();
    L.maser = c->d_maser.as<int>();
    L.cmask = cmask ? c->d_cmask.as<unsigned char>() : nullptr;
    // tile mapping: autotuned once per batch shape (rays per wave, channel
    // tiling, persistent grid size)
    int dev_cu = 0;
    HIPCHK(c, hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device));
    const int bpc = rl::formal_batch_blocks_per_cu(fp64);
    std::string shape = std::to_string(fp64);
    for (int k = 0; k < nb; k++)
        shape += ":" + std::to_string(hi[k].nrays) + "x" + std::to_string(hi[k].nfr) + "s" +
                 std::to_string(hi[k].skip_mode);
    auto set_tiles = [&](int cw) -> int {
        std::vector<long long> &t0 = c->h_tile0;
        t0.assign(nb + 1, 0);
        for (int k = 0; k < nb; k++) t0[k + 1] = t0[k] + tiles_of(cw, hi[k].nrays, hi[k].nfr);
        L.ntiles = t0[nb];
        return upload(c, c->d_tile0, t0.data(), (size_t)(nb + 1) * 8);
    };
    auto zero_counters = [&]() {
        rl::launch_zero_u64(c->stream, L.ovs, nb);
        rl::launch_zero_u64(c->stream, reinterpret_cast<unsigned long long *>(L.maser),
                            (nb + 1) / 2);
    };
    // cell tables of the batch's lines
    auto cells = [&]() {
        for (size_t q = 0; q < slot_line.size(); q++) {
            const int ln = slot_line[q];
            rl::launch_cells(c->stream, ncell, c->nlev, c->nlines, ln, c->lev_up[ln],
                             c->lev_low[ln], molpg, c->d_rho.as<double>(),
                             c->d_abund.as<double>(), c->d_lw.as<double>(),
                             c->d_pops.as<double>(), c->d_jd.as<double>(),
                             c->d_ad.as<double>(), c->d_cells.as<CellRec>() + q * ncell);
        }
    };
    int cw = 0, nblk = dev_cu * bpc;
    auto tuned = c->tune.find(shape);
    if (c->fixed_cw >= 0) {
        cw = c->fixed_cw;
        nblk = c->fixed_persistent ? dev_cu * bpc : 0;
    } else if (tuned != c->tune.end()) {
        cw = tuned->second.first;
        nblk = tuned->second.second;
    } else if (c->autotune) {
        // trials on the first tiles of the batch (enough to fill the device
        // several times): static grid or persistent grid, x tile mapping
        cells();
        rl::launch_batch_nodes(c->stream, fp64, L);
        float best = 1e30f;
        const int cws[4] = {0, 64, 32, 16};
        const int grids[2] = {0, dev_cu * bpc};
        for (int ci = 0; ci < 4; ci++) {
            if (int e = set_tiles(cws[ci])) return e;
            const long long ntr = std::min<long long>(L.ntiles, 64LL * 4 * dev_cu * bpc);
            for (int gi = 0; gi < 2; gi++) {
                zero_counters();
                HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
                rl::launch_batch_formal(c->stream, fp64, L, cws[ci], grids[gi], ntr);
                HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
                HIPCHK(c, hipEventSynchronize(c->ev[3]));
                float t = 0;
                HIPCHK(c, hipEventElapsedTime(&t, c->ev[2], c->ev[3]));
                // the whole batch at this rate: tiles hold 64 lanes in every
                // mapping, but a channel tiling leaves the lanes past the last
                // channel idle, so the mappings have different tile counts
                t = t / (float)ntr * (float)L.ntiles;
                if (t < best) { best = t; cw = cws[ci]; nblk = grids[gi]; }
            }
        }
        c->tune[shape] = {cw, nblk};
    }
    c->last_cw = cw;
    c->last_blocks = nblk;
    if (int e = set_tiles(cw)) return e;
    HIPCHK(c, hipMemsetAsync(c->d_err.p, 0, 4, c->stream));
    // the device stages; recorded once per (shape, buffers) into a graph
    auto stages = [&](bool timed) -> int {
        zero_counters();
        cells();
        rl::launch_batch_nodes(c->stream, fp64, L);
        if (timed) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
        rl::launch_batch_formal(c->stream, fp64, L, cw, nblk, L.ntiles);
        rl::launch_batch_finish(c->stream, fp64, L);
        if (timed) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
        rl::launch_batch_flux(c->stream, L);
        return 0;
    };
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    // everything the recorded launches take by value
    std::string gkey = shape;
    {
        char buf[64];
        auto add = [&](const void *p) { snprintf(buf, sizeof buf, ":%p", p); gkey += buf; };
        for (const void *p : {c->d_rec.p, c->d_image.p, c->d_partial.p, c->d_flux.p,
                              c->d_rinfo.p, c->d_cells.p, c->d_ovs.p, c->d_maser.p,
                              c->d_counter.p, c->d_items.p, c->d_tile0.p, c->d_chan.p,
                              (void *)L.cmask,
                              c->d_rho.p, c->d_abund.p, c->d_lw.p, c->d_pops.p, c->d_jd.p,
                              c->d_ad.p})
            add(p);
        snprintf(buf, sizeof buf, ":%d:%d:%d:%d:%d:%.17g", cw, nblk, c->nlev, c->nlines, ncell,
                 molpg);
        gkey += buf;
        for (int ln : slot_line) {
            snprintf(buf, sizeof buf, ":%d:%d:%d", ln, c->lev_up[ln], c->lev_low[ln]);
            gkey += buf;
        }
    }
    if (c->use_graph) {
        auto g = c->graphs.find(gkey);
        if (g == c->graphs.end() && c->graphs.size() >= 32) {
            for (auto &q : c->graphs) (void)hipGraphExecDestroy(q.second);
            c->graphs.clear();
        }
        if (g == c->graphs.end()) {
            hipGraph_t graph;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int e = stages(false);
            hipError_t ec = hipStreamEndCapture(c->stream, &graph);
            if (e) return e;
            HIPCHK(c, ec);
            hipGraphExec_t ex;
            HIPCHK(c, hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            g = c->graphs.emplace(gkey, ex).first;
        }
        HIPCHK(c, hipGraphLaunch(g->second, c->stream));
        HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    } else {
        if (int e = stages(true)) return e;
        HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(flux + fo, c->d_flux.p, (size_t)flux0 * 8, hipMemcpyDeviceToHost,
                             c->stream));
    if (image)
        HIPCHK(c, hipMemcpyAsync(image + io, c->d_image.p, (size_t)img0 * 8,
                                 hipMemcpyDeviceToHost, c->stream));
    if (cmask)
        HIPCHK(c, hipMemcpyAsync(cmask + io, c->d_cmask.p, (size_t)img0, hipMemcpyDeviceToHost,
                                 c->stream));
    std::vector<unsigned long long> ovs(nb);
    std::vector<int> maser(nb + 1);
    HIPCHK(c, hipMemcpyAsync(ovs.data(), L.ovs, (size_t)nb * 8, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(maser.data(), L.maser, (size_t)nb * 4, hipMemcpyDeviceToHost,
                             c->stream));
    if (int e = check_err(c)) return e;
    unsigned long long ovs_last = 0;
    for (int k = 0; k < nb; k++) {
        if (warn)
            warn[i0 + k] = (maser[k] ? RL_WARN_MASER : 0u) | (ovs[k] ? RL_WARN_OVERSHOOT : 0u);
        ovs_last = ovs[k];
    }
    c->counts[2] = (long long)ovs_last;
    float t14 = 0, t12 = 0, t23 = 0, t34 = 0;
    (void)hipEventElapsedTime(&t14, c->ev[1], c->ev[4]);
    if (c->use_graph) {
        c->t_ms[2] += t14;           // nodes, formal and flux run as one graph
    } else {
        (void)hipEventElapsedTime(&t12, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&t23, c->ev[2], c->ev[3]);
        (void)hipEventElapsedTime(&t34, c->ev[3], c->ev[4]);
        c->t_ms[1] += t12; c->t_ms[2] += t23; c->t_ms[3] += t34;
    }
    c->t_ms[4] += t14;
    return 0;
}

extern "C" int rl_trace_lines_forward(rl_ctx *c, int nitems, const rl_item *items,
                                      double *flux, double *image, unsigned *warn) {
    return rl_trace_lines_forward_ex(c, nitems, items, flux, image, nullptr, warn);
}

extern "C" int rl_trace_lines_forward_ex(rl_ctx *c, int nitems, const rl_item *items,
                                         double *flux, double *image, unsigned char *cmask,
                                         unsigned *warn) {
    if (!c || nitems < 0 || (nitems && (!items || !flux)))
        return fail(c, RL_E_ARG, "rl_trace_lines_forward: bad args");
    if (!c->have_medium || !c->nlev || !c->nlines || !c->have_dust || c->incl.empty())
        return fail(c, 13, "inputs incomplete: grid, medium, levels, lines, dust, rays");
    if (c->outer_bc == 3 && c->isrf.empty())    // read_interstellfield, star.F:697-703
        return fail(c, 13, "outer boundary type 3 without the interstellar field");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_err.ensure(4));
    for (int i = 0; i < 5; i++) c->t_ms[i] = 0.0;
    for (int i = 0; i < 4; i++) c->counts[i] = 0;
    for (int ii = 0; ii < nitems; ii++) {
        const rl_item &it = items[ii];
        if (it.line < 0 || it.line >= c->nlines || it.incl < 0 ||
            it.incl >= (int)c->incl.size() || it.nfr < 2 || !(it.vmax_kms > 0.0) ||
            (it.skip_mode != RL_SKIP_REFERENCE && it.skip_mode != RL_SKIP_NONE))
            return fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", ii);
        if (c->incl[it.incl].rect)
            return fail(c, 13, "item %d: a rectangular-image ray set: use rl_trace_lines_posvel",
                        ii);
    }
    // node counts of every inclination used: an inclination whose nodes
    // surely fit (at the per-ray maximum, within half the budget beside what
    // is resident) is built at once, the others only counted (k_cross) and
    // built per batch; a ray set not used by a batch may be evicted to make
    // room for it
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    {
        std::vector<int> now, later;
        size_t room = resident_traj(c);
        for (int ii = 0; ii < nitems; ii++) {
            const int in = items[ii].incl;
            const RaySet &rs = c->incl[in];
            if (rs.counted || std::find(now.begin(), now.end(), in) != now.end() ||
                std::find(later.begin(), later.end(), in) != later.end())
                continue;
            const size_t most = (size_t)rs.nrays * traj_nmax(c) * (sizeof(GeoNode) + 12);
            if ((room + most) * 2 <= ctx_budget(c) && (int)now.size() < rl::CROSS_SETS_MAX) {
                now.push_back(in);
                room += most;
            } else {
                later.push_back(in);
            }
        }
        if (!now.empty()) {
            if (int e = cross_sets(c, now)) return e;
            for (int in : now)
                if (int e = build_traj_keep(c, in, now)) return e;
        }
        for (int in : later)
            if (int e = count_traj(c, in)) return e;
        release_cross_scratch(c);
    }
    std::vector<std::pair<int, int>> cuts;
    c->last_work = cut_batches(c, nitems, items, cuts);
    release_adjoint_work(c);
    HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[5]));
    float tt = 0;
    (void)hipEventElapsedTime(&tt, c->ev[0], c->ev[5]);
    c->t_ms[0] = tt;
    c->t_ms[4] = tt;
    c->last_batches = 0;
    size_t fo = 0, io = 0;
    for (const auto &cut : cuts) {
        const int i0 = cut.first, i1 = cut.second;
        // this batch's trajectories resident (others evicted if needed)
        std::vector<int> incs;
        size_t ibytes = 0;
        for (int k = i0; k < i1; k++) {
            const int in = items[k].incl;
            ibytes += item_bytes(c, items[k], c->incl[in]);
            if (std::find(incs.begin(), incs.end(), in) == incs.end()) incs.push_back(in);
        }
        for (int in : incs)
            if (!c->incl[in].traj_ready) {
                const size_t want = traj_bytes(c, c->incl[in]), have = held_bytes(c->incl[in]);
                ibytes += want > have ? want - have : 0;
            }
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        if (int e = make_room(c, ibytes, incs)) return e;
        for (int in : incs)
            if (int e = build_traj_keep(c, in, incs)) return e;
        release_cross_scratch(c);
        HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev[5]));
        (void)hipEventElapsedTime(&tt, c->ev[0], c->ev[5]);
        c->t_ms[0] += tt;
        c->t_ms[4] += tt;
        if (int e = run_batch(c, items, i0, i1, flux, image, cmask, warn, fo, io)) return e;
        c->last_batches++;
        for (int k = i0; k < i1; k++) {
            fo += items[k].nfr;
            io += (size_t)c->incl[items[k].incl].nrays * items[k].nfr;
        }
    }
    return 0;
}

// rl_trace_lines_forward without device allocation: every work buffer of
// the call (cell tables, node records, ray windows, channel tables, ring
// partial sums, counters, item descriptors, and the image when the caller
// does not take it) is carved from the library's fixed workspace, and the
// flux (and image) are written to the caller's device memory.  The items
// run as one batch through the forward's kernels; the trajectories of
// their inclinations must already be resident (built by an earlier
// allocating call), and no graph or tile trial is made (a tile mapping the
// autotuner chose earlier for the same batch shape is reused).
extern "C" int rl_trace_lines_forward_noalloc(rl_ctx *c, int nitems, const rl_item *items,
                                              double *flux_dev, double *image_dev,
                                              unsigned *warn) {
    if (!c || nitems < 1 || !items || !flux_dev)
        return fail(c, RL_E_ARG, "rl_trace_lines_forward_noalloc: bad args");
    if (!c->scratch)
        return fail(c, RL_E_HIP, "rl_trace_lines_forward_noalloc: no workspace on device %d (%s)",
                    c->device, c->scratch_err.c_str());
    if (!c->have_medium || !c->nlev || !c->nlines || !c->have_dust || c->incl.empty())
        return fail(c, 13, "inputs incomplete: grid, medium, levels, lines, dust, rays");
    if (c->outer_bc == 3 && c->isrf.empty())    // read_interstellfield, star.F:697-703
        return fail(c, 13, "outer boundary type 3 without the interstellar field");
    const int fp64 = c->precision == RL_FP64;
    const size_t rsz = fp64 ? sizeof(rl::NodeRec<double>) : sizeof(rl::NodeLine);
    const int ncell = c->nr * c->nh;
    const double molpg = 1.0 / (c->umass_av * 1.6726e-24);
    const int nb = nitems;
    std::vector<rl::ItemDev> hi(nb);
    std::vector<int> slot_line;
    size_t nchan = 0;
    for (int k = 0; k < nb; k++) {
        const rl_item &it = items[k];
        if (it.line < 0 || it.line >= c->nlines || it.incl < 0 ||
            it.incl >= (int)c->incl.size() || it.nfr < 2 || !(it.vmax_kms > 0.0) ||
            (it.skip_mode != RL_SKIP_REFERENCE && it.skip_mode != RL_SKIP_NONE))
            return fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", k);
        const RaySet &rs = c->incl[it.incl];
        if (rs.rect)
            return fail(c, 13, "item %d: a rectangular-image ray set: use rl_trace_lines_posvel", k);
        if (!rs.traj_ready || (!fp64 && !(rs.seg_ready && rs.seg_lw_gen == c->lw_gen)))
            return fail(c, 13,
                        "item %d: the trajectories of inclination %d are not resident (an "
                        "allocating trace of that inclination builds them)", k, it.incl);
        nchan += 4 * (size_t)it.nfr;
    }
    std::vector<double> chan(nchan, 0.0);
    long long rec0 = 0, img0 = 0, chan0 = 0, part0 = 0, flux0 = 0;
    int rinfo0 = 0, max_rays = 0, max_nfr = 0, max_rings = 0, any_skip = 0;
    for (int k = 0; k < nb; k++) {
        const rl_item &it = items[k];
        const RaySet &rs = c->incl[it.incl];
        rl::ItemDev &d = hi[k];
        if (int e = item_consts(c, it, rs, d.lc, chan.data() + chan0)) return e;
        int slot = -1;
        for (size_t q = 0; q < slot_line.size(); q++)
            if (slot_line[q] == it.line) slot = (int)q;
        if (slot < 0) { slot = (int)slot_line.size(); slot_line.push_back(it.line); }
        d.nrays = rs.nrays; d.nfr = it.nfr; d.nrings = rs.nrings; d.nphi = rs.nphi;
        d.skip_mode = it.skip_mode; d.cslot = slot;
        d.off = rs.d_off.as<long long>();
        d.geo = rs.d_nodes.as<GeoNode>();
        d.hot_off = (long long)(uintptr_t)rs.d_hot.p;     // made relative below
        d.aux_off = (long long)(uintptr_t)rs.d_aux.p;
        d.rec0 = rec0; d.img0 = img0; d.chan0 = chan0; d.part0 = part0; d.flux0 = flux0;
        d.rinfo0 = rinfo0;
        d.aks2 = 2.0 * (c->aksmax / 2.99792458e5);
        d.ring_w = rs.d_ringw.as<double>();
        d.center_w = rs.center_w;
        rec0 += rs.nnodes; img0 += (long long)rs.nrays * it.nfr; chan0 += 4 * it.nfr;
        part0 += (long long)rs.nrings * it.nfr; flux0 += it.nfr; rinfo0 += rs.nrays;
        max_rays = std::max(max_rays, rs.nrays);
        max_nfr = std::max(max_nfr, it.nfr);
        max_rings = std::max(max_rings, rs.nrings);
        any_skip |= it.skip_mode == RL_SKIP_REFERENCE;
    }
    // tile mapping: the autotuner's earlier choice for this batch shape, else
    // ray-major tiles on the persistent grid
    int dev_cu = 0;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device));
    std::string shape = std::to_string(fp64);
    for (int k = 0; k < nb; k++)
        shape += ":" + std::to_string(hi[k].nrays) + "x" + std::to_string(hi[k].nfr) + "s" +
                 std::to_string(hi[k].skip_mode);
    int cw = 0, nblk = dev_cu * rl::formal_batch_blocks_per_cu(fp64);
    if (c->fixed_cw >= 0) {
        cw = c->fixed_cw;
        if (!c->fixed_persistent) nblk = 0;
    } else {
        auto tuned = c->tune.find(shape);
        if (tuned != c->tune.end()) { cw = tuned->second.first; nblk = tuned->second.second; }
    }
    std::vector<long long> t0(nb + 1, 0);
    for (int k = 0; k < nb; k++) t0[k + 1] = t0[k] + tiles_of(cw, hi[k].nrays, hi[k].nfr);
    // the workspace layout, every part 256-byte aligned
    size_t used = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = used;
        used += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_cells = carve(slot_line.size() * ncell * sizeof(CellRec));
    const size_t o_rec = carve((size_t)rec0 * rsz);
    const size_t o_rinfo = carve((size_t)rinfo0 * sizeof(RayInfo));
    const size_t o_chan = carve(nchan * 8);
    const size_t o_img = image_dev ? 0 : carve((size_t)img0 * 8);
    const size_t o_part = carve((size_t)part0 * 8);
    const size_t o_ovs = carve((size_t)nb * 8);
    const size_t o_maser = carve((size_t)((nb + 1) / 2) * 8);
    const size_t o_counter = carve(8);
    const size_t o_items = carve((size_t)nb * sizeof(rl::ItemDev));
    const size_t o_tile0 = carve((size_t)(nb + 1) * 8);
    if (used > rl::SCRATCH_BYTES)
        return fail(c, 13,
                    "rl_trace_lines_forward_noalloc: the items need %zu bytes of workspace, "
                    "more than its %zu (node records %lld x %zu B)", used, rl::SCRATCH_BYTES,
                    rec0, rsz);
    unsigned char *ws = static_cast<unsigned char *>(c->scratch);
    for (auto &d : hi) {      // segment records relative to the record buffer
        d.hot_off -= (long long)(uintptr_t)(ws + o_rec);
        d.aux_off -= (long long)(uintptr_t)(ws + o_rec);
    }
    rl::BatchLaunch L;
    L.nitems = nb; L.max_rays = max_rays; L.max_nfr = max_nfr; L.max_rings = max_rings;
    L.ncell = ncell; L.any_skip = any_skip;
    L.items = reinterpret_cast<const rl::ItemDev *>(ws + o_items);
    L.cells = reinterpret_cast<const CellRec *>(ws + o_cells);
    L.rec = ws + o_rec;
    L.rinfo = reinterpret_cast<RayInfo *>(ws + o_rinfo);
    L.chan = reinterpret_cast<const double *>(ws + o_chan);
    L.image = image_dev ? image_dev : reinterpret_cast<double *>(ws + o_img);
    L.partial = reinterpret_cast<double *>(ws + o_part);
    L.flux = flux_dev;
    L.cmask = nullptr;
    L.d2 = (1.0 * 3.08572e18) * (1.0 * 3.08572e18);
    L.tile0 = reinterpret_cast<const long long *>(ws + o_tile0);
    L.ntiles = t0[nb];
    L.ovs = reinterpret_cast<unsigned long long *>(ws + o_ovs);
    L.maser = reinterpret_cast<int *>(ws + o_maser);
    L.counter = reinterpret_cast<unsigned long long *>(ws + o_counter);
    // host tables in; counters cleared by a kernel (the workspace keeps the
    // previous call's contents)
    HIPCHK(c, hipMemcpyAsync(ws + o_chan, chan.data(), nchan * 8, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(ws + o_items, hi.data(), (size_t)nb * sizeof(rl::ItemDev),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(ws + o_tile0, t0.data(), (size_t)(nb + 1) * 8,
                             hipMemcpyHostToDevice, c->stream));
    rl::launch_zero_u64(c->stream, L.ovs, nb);
    rl::launch_zero_u64(c->stream, reinterpret_cast<unsigned long long *>(L.maser), (nb + 1) / 2);
    for (size_t q = 0; q < slot_line.size(); q++) {
        const int ln = slot_line[q];
        rl::launch_cells(c->stream, ncell, c->nlev, c->nlines, ln, c->lev_up[ln], c->lev_low[ln],
                         molpg, c->d_rho.as<double>(), c->d_abund.as<double>(),
                         c->d_lw.as<double>(), c->d_pops.as<double>(), c->d_jd.as<double>(),
                         c->d_ad.as<double>(),
                         reinterpret_cast<CellRec *>(ws + o_cells) + q * ncell);
    }
    c->cells_line = -1;
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    rl::launch_batch_nodes(c->stream, fp64, L);
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    rl::launch_batch_formal(c->stream, fp64, L, cw, nblk, L.ntiles);
    rl::launch_batch_finish(c->stream, fp64, L);
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    rl::launch_batch_flux(c->stream, L);
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    HIPCHK(c, hipGetLastError());
    std::vector<unsigned long long> ovs(nb);
    std::vector<int> maser(nb + 1, 0);
    HIPCHK(c, hipMemcpyAsync(ovs.data(), L.ovs, (size_t)nb * 8, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(maser.data(), L.maser, (size_t)nb * 4, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < 5; i++) c->t_ms[i] = 0.0;
    for (int i = 0; i < 4; i++) c->counts[i] = 0;
    float t12 = 0, t23 = 0, t34 = 0;
    (void)hipEventElapsedTime(&t12, c->ev[1], c->ev[2]);
    (void)hipEventElapsedTime(&t23, c->ev[2], c->ev[3]);
    (void)hipEventElapsedTime(&t34, c->ev[3], c->ev[4]);
    c->t_ms[1] = t12; c->t_ms[2] = t23; c->t_ms[3] = t34; c->t_ms[4] = t12 + t23 + t34;
    for (int k = 0; k < nb; k++) {
        c->counts[0] += hi[k].nrays;
        c->counts[1] += c->incl[items[k].incl].nnodes;
        c->counts[3] += (long long)hi[k].nrays * hi[k].nfr;
        if (warn)
            warn[k] = (maser[k] ? RL_WARN_MASER : 0u) | (ovs[k] ? RL_WARN_OVERSHOOT : 0u);
    }
    c->counts[2] = (long long)ovs[nb - 1];
    c->last_cw = cw;
    c->last_blocks = nblk;
    c->last_batches = 1;
    return 0;
}

// Circular image -> rectangular grid (the interpolation of imcir_conv,
// PRO/imcir_conv.pro:161-181): out[ch][p] = sum_k w[p][k] image[vtx[p][k]][ch],
// the triangulation weights formed by the caller.  Every vertex index is
// checked against nrays before the launch.
extern "C" int rl_resample_image(rl_ctx *c, int nrays, int nfr, const double *image, int npix,
                                 int nv, const int *vtx, const double *w, double *out) {
    if (!c || nrays < 1 || nfr < 1 || npix < 0 || nv < 1 || !image || (npix && (!vtx || !w || !out)))
        return fail(c, RL_E_ARG, "rl_resample_image: bad args");
    const size_t nvt = (size_t)npix * nv;
    for (size_t i = 0; i < nvt; i++)
        if (vtx[i] >= nrays || vtx[i] < -1)
            return fail(c, RL_E_ARG, "rl_resample_image: vertex %d out of range", vtx[i]);
    if (npix == 0) return 0;
    HIPCHK(c, hipSetDevice(c->device));
    DevBuf d_img, d_vtx, d_w, d_out;
    if (int e = upload(c, d_img, image, (size_t)nrays * nfr * 8)) return e;
    if (int e = upload(c, d_vtx, vtx, nvt * 4)) return e;
    if (int e = upload(c, d_w, w, nvt * 8)) return e;
    HIPCHK(c, d_out.ensure((size_t)npix * nfr * 8));
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    rl::launch_rect_gather(c->stream, npix, nfr, nv, d_vtx.as<int>(), d_w.as<double>(),
                           d_img.as<double>(), d_out.as<double>());
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    HIPCHK(c, hipMemcpyAsync(out, d_out.p, (size_t)npix * nfr * 8, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));        // the buffers are freed on return
    float t = 0;
    (void)hipEventElapsedTime(&t, c->ev[1], c->ev[4]);
    for (int i = 0; i < 5; i++) c->t_ms[i] = 0.0;
    c->t_ms[2] = t;
    c->t_ms[4] = t;
    return 0;
}

// Position-velocity cubes (calc_write_line_posvel, telescope.F:1828-2031;
// make_freq_image_rectang, telescope.F:2061-2215): per item, the
// intensity and the optical depth of every pixel ray at every channel,
// every channel integrated (no NONREDUNDANT skipping: the rectangular
// imager has none), pixels outside 0.999 R_max black with tau 0, and the
// unresolved star added to the four central pixels.  image / tau are
// [item][iy][ix][channel].
extern "C" int rl_trace_lines_posvel(rl_ctx *c, int nitems, const rl_item *items, double *image,
                                     double *tau, unsigned *warn) {
    if (!c || nitems < 0 || (nitems && (!items || !image)))
        return fail(c, RL_E_ARG, "rl_trace_lines_posvel: bad args");
    if (!c->have_medium || !c->nlev || !c->nlines || !c->have_dust || c->incl.empty())
        return fail(c, 13, "inputs incomplete: grid, medium, levels, lines, dust, rays");
    if (c->outer_bc == 3 && c->isrf.empty())    // read_interstellfield, star.F:697-703
        return fail(c, 13, "outer boundary type 3 without the interstellar field");
    for (int ii = 0; ii < nitems; ii++) {
        const rl_item &it = items[ii];
        if (it.incl < 0 || it.incl >= (int)c->incl.size() || !c->incl[it.incl].rect)
            return fail(c, 13, "item %d: not a rectangular-image ray set "
                               "(rl_setup_rays_rectang)", ii);
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_err.ensure(4));
    for (int i = 0; i < 5; i++) c->t_ms[i] = 0.0;
    for (int i = 0; i < 4; i++) c->counts[i] = 0;
    release_adjoint_work(c);
    const bool fp64 = c->precision == RL_FP64;
    const size_t rsz = fp64 ? sizeof(rl::NodeRec<double>) : sizeof(rl::FastRec);
    size_t io = 0;
    std::vector<double> cen;                      // the added star's ray
    for (int ii = 0; ii < nitems; ii++) {
        rl_item it = items[ii];
        it.skip_mode = RL_SKIP_NONE;
        ItemSetup s;
        if (int e = setup_item(c, it, ii, s, c->ev[0], c->ev[1])) return e;
        release_cross_scratch(c);
        RaySet &rs = *s.rs;
        const int nfr = it.nfr, npix = rs.nx * rs.ny;
        const size_t ncv = (size_t)rs.nrays * nfr;
        const size_t npv = (size_t)npix * nfr, nextra = (size_t)(rs.nrays - npix) * nfr;
        HIPCHK(c, c->d_rec.ensure((size_t)rs.nnodes * rsz));
        HIPCHK(c, c->d_rinfo.ensure((size_t)rs.nrays * sizeof(RayInfo)));
        HIPCHK(c, c->d_image.ensure((npv + nextra + 1) * 8));
        HIPCHK(c, c->d_partial.ensure(npv * 8 + 8));
        HIPCHK(c, c->d_ovs.ensure(16));
        unsigned long long *ovs = c->d_ovs.as<unsigned long long>();
        rl::launch_zero_u64(c->stream, ovs, 2);
        rl::PosvelLaunch L;
        L.nrays = rs.nrays; L.nfr = nfr; L.nx = rs.nx; L.ny = rs.ny;
        L.off = rs.d_off.as<long long>();
        L.geo = rs.d_nodes.as<GeoNode>();
        L.cells = c->d_cells.as<CellRec>();
        L.rinfo = c->d_rinfo.as<RayInfo>();
        L.aks2 = 2.0 * (c->aksmax / 2.99792458e5);
        L.rec = c->d_rec.p;
        L.dnu = c->d_chan.as<double>();
        L.velo = L.dnu + nfr;
        L.istar = L.dnu + 2 * nfr;
        L.i0 = L.dnu + 3 * nfr;
        L.lc = s.lc;
        L.image = c->d_image.as<double>();
        L.extra = L.image + npv;
        L.tau = c->d_partial.as<double>();
        L.ovs = ovs;
        L.maser = reinterpret_cast<int *>(ovs + 1);
        if (fp64) rl::launch_posvel<double>(c->stream, L);
        else rl::launch_posvel<float>(c->stream, L);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
        double *oi = image + io, *ot = tau ? tau + io : nullptr;
        cen.assign(nextra ? nextra : 1, 0.0);
        unsigned long long flags[2] = {0, 0};
        HIPCHK(c, hipMemcpyAsync(oi, L.image, npv * 8, hipMemcpyDeviceToHost, c->stream));
        if (ot) HIPCHK(c, hipMemcpyAsync(ot, L.tau, npv * 8, hipMemcpyDeviceToHost, c->stream));
        if (nextra)
            HIPCHK(c, hipMemcpyAsync(cen.data(), L.extra, nextra * 8, hipMemcpyDeviceToHost,
                                     c->stream));
        HIPCHK(c, hipMemcpyAsync(flags, ovs, 16, hipMemcpyDeviceToHost, c->stream));
        if (int e = check_err(c)) return e;             // synchronises the stream
        if (rs.starunres) {                               // telescope.F:2153-2200
            const double srat = (3.14159265 * c->rstar * c->rstar) / (4.0 * rs.spx * rs.spy);
            const double srat1 = 1.0 - srat;
            const int ixc = rs.nx / 2 - 1, iyc = rs.ny / 2 - 1;
            for (int k = 0; k < nfr; k++) {
                const double dummy = cen[k] * srat;
                for (int dy = 0; dy < 2; dy++)
                    for (int dx = 0; dx < 2; dx++) {
                        double &p = oi[((size_t)(iyc + dy) * rs.nx + ixc + dx) * nfr + k];
                        p = dummy + srat1 * p;
                    }
            }
        }
        if (warn)
            warn[ii] = ((int)flags[1] ? RL_WARN_MASER : 0u) | (flags[0] ? RL_WARN_OVERSHOOT : 0u);
        float t01 = 0, t14 = 0;
        (void)hipEventElapsedTime(&t01, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&t14, c->ev[1], c->ev[4]);
        c->t_ms[0] += t01;
        c->t_ms[2] += t14;
        c->t_ms[4] += t01 + t14;
        c->counts[0] += rs.nrays;
        c->counts[1] += rs.nnodes;
        c->counts[2] = (long long)flags[0];
        c->counts[3] += (long long)ncv;
        io += (size_t)npix * nfr;
    }
    return 0;
}

extern "C" int rl_set_memory_budget(rl_ctx *c, double bytes) {
    if (!c || bytes < 0) return RL_E_ARG;
    c->mem_budget = (size_t)bytes;
    return 0;
}

extern "C" int rl_set_option(rl_ctx *c, int option, int value) {
    if (!c) return RL_E_ARG;
    switch (option) {
        case RL_OPT_GRAPH: c->use_graph = value != 0; return 0;
        case RL_OPT_AUTOTUNE: c->autotune = value != 0; if (!value) c->tune.clear(); return 0;
        case RL_OPT_TILE:
            if (value != 0 && value != 16 && value != 32 && value != 64)
                return fail(c, RL_E_ARG, "rl_set_option: tile mapping must be 0, 16, 32 or 64");
            c->tune.clear();
            c->autotune = false;
            c->fixed_cw = value;
            return 0;
        case RL_OPT_PERSISTENT:
            c->tune.clear();
            c->fixed_persistent = value != 0;
            return 0;
        case RL_OPT_CKPT:
            if (value < 0 || value > rl::adjoint_block_len() || (value & (value - 1)))
                return fail(c, RL_E_ARG, "rl_set_option: checkpoint interval must be 0 or a "
                                         "power of two up to %d", rl::adjoint_block_len());
            c->fixed_kblk = value;
            return 0;
        default: return fail(c, RL_E_ARG, "rl_set_option: unknown option %d", option);
    }
}

extern "C" int rl_last_batch_info(const rl_ctx *c, int *batches, int *cw, int *blocks) {
    if (!c) return RL_E_ARG;
    if (batches) *batches = c->last_batches;
    if (cw) *cw = c->last_cw;
    if (blocks) *blocks = c->last_blocks;
    return 0;
}

static int zero_dev(rl_ctx *c, DevBuf &b, size_t bytes) {
    HIPCHK(c, b.ensure(bytes));
    HIPCHK(c, hipMemsetAsync(b.p, 0, bytes, c->stream));
    return 0;
}

// A chunk of an item's rays: rays [r0, r1), nodes [n0, n1).
struct AdjChunk {
    int r0, r1;
    long long n0, n1;
};

// checkpoint interval and ray chunks of one adjoint item
struct AdjPlan {
    int K = 1;
    std::vector<AdjChunk> chunks;
    size_t work = 0;                 // bytes of the largest chunk's buffers + the image
};

// device bytes of a chunk's buffers at interval K: node records, the I
// checkpoints, node gradients (scaled and true), the checkpoint CSR and,
// for a chunk of a split item, its own stencil transpose with the sort's
// scratch
static size_t chunk_bytes(const rl_ctx *c, const RaySet &rs, int nfr, int K, int r0, int r1,
                          bool own_transpose) {
    const int fp64 = c->precision == RL_FP64;
    const size_t tsz = fp64 ? 8 : 4;
    const size_t nn = (size_t)(rs.off[r1] - rs.off[r0]), nr = (size_t)(r1 - r0);
    size_t b = nn * rl::fast_rec_size(fp64);
    b += (nn / K + nr) * (size_t)nfr * tsz;
    b += nn * 5 * 8 * 2;
    b += (nr + 1) * 8;
    if (own_transpose) b += nn * 4 * 24 + ((size_t)c->nr * c->nh + 1) * 4 + (nn * 4 * 8 >> 3);
    return b;
}

// The plan of an item: the smallest interval K (or the fixed one) at which
// the whole item fits the budget beside its trajectories, their stencil
// transpose and its image; else the largest interval (or the fixed one)
// with the rays cut into chunks of about equal node counts that fit beside
// the trajectories and the image (at least 2, at most 64 chunks; at most
// 2^28 nodes each, within the transpose's 32-bit entry index).
static int plan_adjoint(rl_ctx *c, const RaySet &rs, int nfr, bool tape, AdjPlan &p) {
    const int KMAX = rl::adjoint_block_len();
    const size_t budget = ctx_budget(c);
    const size_t nn = (size_t)rs.nnodes;
    const size_t base = nn * sizeof(GeoNode) + (size_t)(rs.nrays + 1) * 8 +
                        (size_t)rs.nrays * nfr * 8 + (size_t)rs.nrings * nfr * 8;
    const size_t transpose = nn * 4 * 12 + ((size_t)c->nr * c->nh + 1) * 4;
    p.chunks.clear();
    const bool fits_index = 4 * nn < ((size_t)1 << 31);
    for (int K = c->fixed_kblk ? c->fixed_kblk : 1; K <= KMAX; K *= 2) {
        const size_t b = chunk_bytes(c, rs, nfr, K, 0, rs.nrays, false);
        if (base + transpose + b <= budget && fits_index) {
            p.K = K;
            p.chunks.push_back({0, rs.nrays, 0, rs.nnodes});
            p.work = b + (size_t)rs.nrays * nfr * 8;
            return 0;
        }
        if (c->fixed_kblk) break;
    }
    if (tape)
        return fail(c, 13, "adjoint tape of %lld nodes x %d channels does not fit the memory "
                           "budget (%.3g GB); rl_trace_lines_adjoint recomputes in ray chunks",
                    rs.nnodes, nfr, budget * 1e-9);
    p.K = c->fixed_kblk ? c->fixed_kblk : KMAX;
    const size_t avail = budget > base ? budget - base : 0;
    const size_t whole = chunk_bytes(c, rs, nfr, p.K, 0, rs.nrays, true);
    long long nch = avail ? (long long)((whole + avail - 1) / avail) : 64;
    nch = std::max(nch, (long long)(rs.nnodes >> 28) + 1);      // <= 2^28 nodes a chunk
    nch = std::min(std::max(nch, 2ll), 64ll);
    nch = std::min(nch, (long long)rs.nrays);
    // chunk j ends at the first ray boundary at or past j/nch of the nodes
    int r0 = 0;
    p.work = 0;
    for (long long j = 1; j <= nch && r0 < rs.nrays; j++) {
        int r1 = rs.nrays;
        if (j < nch) {
            const long long target = rs.nnodes * j / nch;
            r1 = (int)(std::lower_bound(rs.off.begin() + r0 + 1, rs.off.end(), target) -
                       rs.off.begin());
            r1 = std::min(std::max(r1, r0 + 1), rs.nrays);
        }
        p.chunks.push_back({r0, r1, rs.off[r0], rs.off[r1]});
        p.work = std::max(p.work, chunk_bytes(c, rs, nfr, p.K, r0, r1, true));
        r0 = r1;
    }
    p.work += (size_t)rs.nrays * nfr * 8;
    return 0;
}

static void fill_adj(rl_ctx *c, const rl_item &it, const ItemSetup &su, AdjItemBufs &ab, int K,
                     const AdjChunk &ch, rl::AdjLaunch &a) {
    const RaySet &rs = *su.rs;
    const int nfr = it.nfr;
    a.nrays = ch.r1 - ch.r0; a.nfr = nfr; a.nrings = rs.nrings; a.nphi = rs.nphi;
    a.nnodes = ch.n1 - ch.n0;
    a.ray0 = ch.r0;
    a.node0 = ch.n0;
    a.ncell = c->nr * c->nh; a.nlev = c->nlev; a.nlines = c->nlines; a.line = it.line;
    a.lev_up = c->lev_up[it.line]; a.lev_low = c->lev_low[it.line];
    a.molpg = 1.0 / (c->umass_av * 1.6726e-24);
    a.off = rs.d_off.as<long long>() + ch.r0;
    a.ckoff = ab.ckoff.as<long long>();
    a.geo = rs.d_nodes.as<GeoNode>();
    a.cells = c->d_cells.as<CellRec>();
    a.rec = ab.rec.p;
    a.rinfo = c->d_rinfo.as<RayInfo>();
    a.ckpt = ab.ckpt.p;
    a.kblk = K;
    a.dnu = c->d_chan.as<double>();
    a.velo = a.dnu + nfr;
    a.istar = a.dnu + 2 * nfr;
    a.i0 = a.dnu + 3 * nfr;
    a.lc = su.lc;
    a.aks2 = 2.0 * (c->aksmax / 2.99792458e5);
    a.image = c->d_image.as<double>() + (size_t)ch.r0 * nfr;
    a.ring_w = rs.d_ringw.as<double>();
    a.center_w = rs.center_w;
    a.nchunks = rs.nrings;
    a.partial = c->d_partial.as<double>();
    a.flux = c->d_flux.as<double>();
    a.d2 = (1.0 * 3.08572e18) * (1.0 * 3.08572e18);
    a.gflux = c->d_gflux.as<double>();
    a.gnode = c->d_gnode.as<double>();
    a.gtrue = c->d_gcell.as<double>();
    const bool whole = ch.r0 == 0 && ch.r1 == rs.nrays;
    a.st_off = whole ? rs.d_st_off.as<int>() : c->d_stc_off.as<int>();
    a.st_vals = whole ? rs.d_st_vals.as<int>() : c->d_stc_vals.as<int>();
    a.st_w = whole ? rs.d_st_w.as<double>() : c->d_stc_w.as<double>();
    a.rho = c->d_rho.as<double>();
    a.abund = c->d_abund.as<double>();
    a.pops = c->d_pops.as<double>();
    a.g_rho = c->d_grho.as<double>();
    a.g_abund = c->d_gabund.as<double>();
    a.g_lw = c->d_glw.as<double>();
    a.g_pops = c->d_gpops.as<double>();
    a.g_jd = c->d_gjd.as<double>();
    a.g_ad = c->d_gad.as<double>();
}

// the item's trajectories resident (other inclinations evicted so that
// they and work bytes of buffers fit the budget), its line tables, and the
// image / flux buffers
static int adj_setup(rl_ctx *c, const rl_item &it, int ii, ItemSetup &su, size_t work) {
    if (it.incl < 0 || it.incl >= (int)c->incl.size())
        return fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", ii);
    const std::vector<int> keep{it.incl};
    if (int e = make_room(c, work, keep)) return e;
    if (int e = build_traj_keep(c, it.incl, keep)) return e;
    release_cross_scratch(c);
    if (int e = setup_item(c, it, ii, su, c->ev[0], c->ev[1])) return e;
    const RaySet &rs = *su.rs;
    HIPCHK(c, c->d_rinfo.ensure((size_t)rs.nrays * sizeof(RayInfo)));
    HIPCHK(c, c->d_image.ensure((size_t)rs.nrays * it.nfr * 8));
    HIPCHK(c, c->d_partial.ensure((size_t)rs.nrings * it.nfr * 8));
    HIPCHK(c, c->d_flux.ensure((size_t)it.nfr * 8));
    return 0;
}

// the checkpointed forward of one chunk: node records, I every K segments
static int adj_forward_chunk(rl_ctx *c, const rl_item &it, const ItemSetup &su, int K,
                             const AdjChunk &ch, AdjItemBufs &ab) {
    const RaySet &rs = *su.rs;
    const int nfr = it.nfr, nr = ch.r1 - ch.r0;
    const size_t tsz = c->precision == RL_FP64 ? 8 : 4;
    std::vector<long long> ckoff(nr + 1, 0);          // ceil((n-1)/K) blocks per ray
    for (int r = 0; r < nr; r++) {
        const long long n = rs.off[ch.r0 + r + 1] - rs.off[ch.r0 + r];
        ckoff[r + 1] = ckoff[r] + (n >= 2 ? (n - 1 + K - 1) / K : 0);
    }
    if (int e = upload(c, ab.ckoff, ckoff.data(), ckoff.size() * 8)) return e;
    HIPCHK(c, ab.ckpt.ensure((size_t)ckoff[nr] * nfr * tsz));
    HIPCHK(c, ab.rec.ensure((size_t)(ch.n1 - ch.n0) * rl::fast_rec_size(c->precision == RL_FP64)));
    rl::AdjLaunch a;
    fill_adj(c, it, su, ab, K, ch, a);
    if (c->precision == RL_FP64) rl::launch_adjoint_fwd<double>(c->stream, a);
    else rl::launch_adjoint_fwd<float>(c->stream, a);
    HIPCHK(c, hipGetLastError());
    c->counts[0] += nr;
    c->counts[1] += ch.n1 - ch.n0;
    return 0;
}

// the spectrum of an item once every chunk is traced
static int adj_flux(rl_ctx *c, const rl_item &it, const ItemSetup &su, double *flux_host) {
    AdjItemBufs none;
    rl::AdjLaunch a;
    fill_adj(c, it, su, none, 1, AdjChunk{0, su.rs->nrays, 0, su.rs->nnodes}, a);
    rl::launch_adjoint_flux(c->stream, a);
    HIPCHK(c, hipGetLastError());
    if (flux_host)
        HIPCHK(c, hipMemcpyAsync(flux_host, c->d_flux.p, (size_t)it.nfr * 8,
                                 hipMemcpyDeviceToHost, c->stream));
    return 0;
}

// the reverse sweep of one chunk (its forward's records and checkpoints),
// accumulating into the context's cell gradients; gflux is on the device
static int adj_reverse_chunk(rl_ctx *c, const rl_item &it, const ItemSetup &su, int K,
                             const AdjChunk &ch, AdjItemBufs &ab) {
    RaySet &rs = *su.rs;
    const long long nn = ch.n1 - ch.n0;
    if (int e = zero_dev(c, c->d_gnode, (size_t)nn * 5 * 8)) return e;
    HIPCHK(c, c->d_gcell.ensure((size_t)nn * 5 * 8));
    if (ch.r0 == 0 && ch.r1 == rs.nrays) {
        if (int e = build_transpose(c, rs)) return e;
    } else {
        if (int e = build_transpose_range(c, rs, ch.n0, ch.n1, c->d_stc_vals, c->d_stc_w,
                                          c->d_stc_off))
            return e;
    }
    rl::AdjLaunch a;
    fill_adj(c, it, su, ab, K, ch, a);
    if (c->precision == RL_FP64) rl::launch_adjoint_rev<double>(c->stream, a);
    else rl::launch_adjoint_rev<float>(c->stream, a);
    HIPCHK(c, hipGetLastError());
    c->counts[3] += (long long)(ch.r1 - ch.r0) * it.nfr;
    return 0;
}

static int adj_check(rl_ctx *c, int nitems, const rl_item *items, const char *who) {
    if (!c->have_medium || !c->nlev || !c->nlines || !c->have_dust || c->incl.empty())
        return fail(c, 13, "inputs incomplete: grid, medium, levels, lines, dust, rays");
    if (c->outer_bc == 3 && c->isrf.empty())    // read_interstellfield, star.F:697-703
        return fail(c, 13, "outer boundary type 3 without the interstellar field");
    for (int ii = 0; ii < nitems; ii++) {
        if (items[ii].skip_mode != RL_SKIP_NONE)
            return fail(c, 13, "%s: item %d: the adjoint integrates every channel (RL_SKIP_NONE)",
                        who, ii);
        const int in = items[ii].incl;
        if (in >= 0 && in < (int)c->incl.size() && c->incl[in].rect)
            return fail(c, 13, "%s: item %d: the adjoint is for circular images", who, ii);
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_err.ensure(4));
    HIPCHK(c, hipMemsetAsync(c->d_err.p, 0, 4, c->stream));
    for (int i = 0; i < 5; i++) c->t_ms[i] = 0.0;
    for (int i = 0; i < 4; i++) c->counts[i] = 0;
    return 0;
}

static bool same_items(const std::vector<rl_item> &a, int n, const rl_item *b) {
    if ((int)a.size() != n) return false;
    for (int i = 0; i < n; i++)
        if (a[i].line != b[i].line || a[i].incl != b[i].incl || a[i].nfr != b[i].nfr ||
            a[i].vmax_kms != b[i].vmax_kms || a[i].skip_mode != b[i].skip_mode)
            return false;
    return true;
}

extern "C" int rl_tape_release(rl_ctx *c) {
    if (!c) return RL_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->tape.clear();
    c->tape_items.clear();
    return 0;
}

extern "C" int rl_trace_lines_forward_tape(rl_ctx *c, int nitems, const rl_item *items,
                                           double *flux) {
    if (!c || nitems < 0 || (nitems && !items))
        return fail(c, RL_E_ARG, "rl_trace_lines_forward_tape: bad args");
    if (int e = adj_check(c, nitems, items, "rl_trace_lines_forward_tape")) return e;
    release_forward_work(c);
    release_adjoint_work(c);
    c->tape_items.clear();                   // invalid until the forward completes
    c->tape.resize(nitems);                  // buffers of a previous tape are reused
    HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    size_t fo = 0;
    for (int ii = 0; ii < nitems; ii++) fo += items[ii].nfr;
    c->tape_flux.assign(fo, 0.0);
    fo = 0;
    c->last_kblk = 0;
    c->last_chunks = 1;
    c->last_work = 0;
    int Kt = 0;
    for (int ii = 0; ii < nitems; ii++) {
        const rl_item &it = items[ii];
        if (it.incl < 0 || it.incl >= (int)c->incl.size())
            return fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", ii);
        if (int e = count_traj(c, it.incl)) return e;
        AdjPlan p;
        if (int e = plan_adjoint(c, c->incl[it.incl], it.nfr, true, p)) return e;
        // one interval for the whole tape: the largest any item needs
        Kt = std::max(Kt, p.K);
        c->last_work = std::max(c->last_work, p.work);
    }
    {
        // every item's records and checkpoints are held at once
        size_t held = 0, fixed = 0;
        for (int ii = 0; ii < nitems; ii++) {
            const RaySet &rs = c->incl[items[ii].incl];
            held += chunk_bytes(c, rs, items[ii].nfr, Kt, 0, rs.nrays, false);
            fixed = std::max(fixed, (size_t)rs.nnodes * (sizeof(GeoNode) + 48) +
                                        (size_t)rs.nrays * items[ii].nfr * 8);
        }
        if (held + fixed > ctx_budget(c))
            return fail(c, 13, "adjoint tape of %d items (%.3g GB) does not fit the memory "
                               "budget (%.3g GB)", nitems, (held + fixed) * 1e-9,
                        ctx_budget(c) * 1e-9);
    }
    for (int ii = 0; ii < nitems; ii++) {
        const rl_item &it = items[ii];
        ItemSetup su;
        if (int e = adj_setup(c, it, ii, su, 0)) return e;
        const AdjChunk whole{0, su.rs->nrays, 0, su.rs->nnodes};
        if (int e = adj_forward_chunk(c, it, su, Kt, whole, c->tape[ii])) return e;
        if (int e = adj_flux(c, it, su, c->tape_flux.data() + fo)) return e;
        fo += it.nfr;
    }
    c->last_kblk = Kt;
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    if (int e = check_err(c)) return e;
    if (flux) std::copy(c->tape_flux.begin(), c->tape_flux.end(), flux);
    float t = 0;
    (void)hipEventElapsedTime(&t, c->ev[5], c->ev[4]);
    c->t_ms[4] = t;
    c->tape_items.assign(items, items + nitems);
    c->tape_gen = c->gen;
    c->tape_kblk = Kt;
    c->last_used_tape = 0;
    return 0;
}

extern "C" int rl_trace_lines_adjoint(rl_ctx *c, int nitems, const rl_item *items,
                                      const double *dL_dflux, double *flux, double *g_rho,
                                      double *g_abund, double *g_lw, double *g_pops,
                                      double *g_jd, double *g_ad) {
    if (!c || nitems < 0 || (nitems && (!items || !dL_dflux)))
        return fail(c, RL_E_ARG, "rl_trace_lines_adjoint: bad args");
    if (int e = adj_check(c, nitems, items, "rl_trace_lines_adjoint")) return e;
    release_forward_work(c);
    const int ncell = c->nr * c->nh;
    if (int e = zero_dev(c, c->d_grho, (size_t)ncell * 8)) return e;
    if (int e = zero_dev(c, c->d_gabund, (size_t)ncell * 8)) return e;
    if (int e = zero_dev(c, c->d_glw, (size_t)ncell * 8)) return e;
    if (int e = zero_dev(c, c->d_gpops, (size_t)ncell * c->nlev * 8)) return e;
    if (int e = zero_dev(c, c->d_gjd, (size_t)ncell * c->nlines * 8)) return e;
    if (int e = zero_dev(c, c->d_gad, (size_t)ncell * c->nlines * 8)) return e;
    HIPCHK(c, hipEventRecord(c->ev[5], c->stream));
    // the tape of the last rl_trace_lines_forward_tape, when it was taken
    // on these items with the current inputs
    const bool use_tape = c->tape_gen == c->gen && same_items(c->tape_items, nitems, items);
    c->last_used_tape = use_tape ? 1 : 0;
    size_t fo = 0;
    if (use_tape) {
        c->last_kblk = c->tape_kblk;
        c->last_chunks = 1;
        if (flux) std::copy(c->tape_flux.begin(), c->tape_flux.end(), flux);
        for (int ii = 0; ii < nitems; ii++) {
            const rl_item &it = items[ii];
            ItemSetup su;
            if (int e = adj_setup(c, it, ii, su, 0)) return e;
            if (int e = upload(c, c->d_gflux, dL_dflux + fo, (size_t)it.nfr * 8)) return e;
            const AdjChunk whole{0, su.rs->nrays, 0, su.rs->nnodes};
            if (int e = adj_reverse_chunk(c, it, su, c->tape_kblk, whole, c->tape[ii])) return e;
            fo += it.nfr;
        }
    } else {
        // items one at a time through the same buffers, each by its own plan
        AdjItemBufs ab;
        ab.rec = std::move(c->d_arec);
        ab.ckoff = std::move(c->d_ckoff);
        ab.ckpt = std::move(c->d_ckpt);
        c->last_kblk = 0;
        c->last_chunks = 0;
        c->last_work = 0;
        int err = 0;
        for (int ii = 0; ii < nitems && !err; ii++) {
            const rl_item &it = items[ii];
            ItemSetup su;
            AdjPlan p;
            err = it.incl < 0 || it.incl >= (int)c->incl.size()
                      ? fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", ii)
                      : count_traj(c, it.incl);
            if (!err) err = plan_adjoint(c, c->incl[it.incl], it.nfr, false, p);
            if (!err && p.chunks.size() > 1 && c->incl[it.incl].st_ready) {
                // the split item's chunks carry their own transposes
                RaySet &rs = c->incl[it.incl];
                HIPCHK(c, hipStreamSynchronize(c->stream));
                rs.d_st_vals = DevBuf();
                rs.d_st_w = DevBuf();
                rs.d_st_off = DevBuf();
                rs.st_ready = false;
            }
            if (!err) err = adj_setup(c, it, ii, su, p.work);
            if (!err) err = upload(c, c->d_gflux, dL_dflux + fo, (size_t)it.nfr * 8);
            for (size_t k = 0; k < p.chunks.size() && !err; k++) {
                err = adj_forward_chunk(c, it, su, p.K, p.chunks[k], ab);
                if (!err) err = adj_reverse_chunk(c, it, su, p.K, p.chunks[k], ab);
            }
            if (!err) err = adj_flux(c, it, su, flux ? flux + fo : nullptr);
            c->last_kblk = std::max(c->last_kblk, p.K);
            c->last_chunks = std::max(c->last_chunks, (int)p.chunks.size());
            c->last_work = std::max(c->last_work, p.work);
            fo += it.nfr;
        }
        c->d_arec = std::move(ab.rec);
        c->d_ckoff = std::move(ab.ckoff);
        c->d_ckpt = std::move(ab.ckpt);
        if (c->last_chunks > 1) {             // per-chunk transposes are not kept
            c->d_stc_vals = DevBuf();
            c->d_stc_w = DevBuf();
            c->d_stc_off = DevBuf();
        }
        if (err) return err;
    }
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    if (int e = check_err(c)) return e;
    float t = 0;
    (void)hipEventElapsedTime(&t, c->ev[5], c->ev[4]);
    c->t_ms[2] = t;
    c->t_ms[4] = t;
    auto down = [&](double *dst, DevBuf &b, size_t n) -> int {
        if (dst) HIPCHK(c, hipMemcpyAsync(dst, b.p, n * 8, hipMemcpyDeviceToHost, c->stream));
        return 0;
    };
    if (int e = down(g_rho, c->d_grho, ncell)) return e;
    if (int e = down(g_abund, c->d_gabund, ncell)) return e;
    if (int e = down(g_lw, c->d_glw, ncell)) return e;
    if (int e = down(g_pops, c->d_gpops, (size_t)ncell * c->nlev)) return e;
    if (int e = down(g_jd, c->d_gjd, (size_t)ncell * c->nlines)) return e;
    if (int e = down(g_ad, c->d_gad, (size_t)ncell * c->nlines)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_last_adjoint_info(const rl_ctx *c, int *ckpt_interval, int *used_tape,
                                    int *chunks) {
    if (!c) return RL_E_ARG;
    if (ckpt_interval) *ckpt_interval = c->last_kblk;
    if (used_tape) *used_tape = c->last_used_tape;
    if (chunks) *chunks = c->last_chunks;
    return 0;
}

extern "C" int rl_plan(rl_ctx *c, int nitems, const rl_item *items, double *out) {
    if (!c || nitems < 0 || (nitems && !items) || !out) return RL_E_ARG;
    if (c->incl.empty()) return fail(c, 13, "inputs incomplete: rays");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->d_err.ensure(4));
    for (int ii = 0; ii < nitems; ii++) {
        const rl_item &it = items[ii];
        if (it.incl < 0 || it.incl >= (int)c->incl.size() || it.nfr < 2)
            return fail(c, 13, "item %d: bad line/incl/nfr/vmax/skip_mode", ii);
        if (int e = count_traj(c, it.incl)) return e;
    }
    release_cross_scratch(c);
    std::vector<std::pair<int, int>> cuts;
    const size_t most = cut_batches(c, nitems, items, cuts);
    int K = 0, nch = 0, tape_ok = 1;
    size_t work = 0;
    const std::string keep_err = c->err;
    for (int ii = 0; ii < nitems; ii++) {
        AdjPlan p;
        if (int e = plan_adjoint(c, c->incl[items[ii].incl], items[ii].nfr, false, p)) return e;
        K = std::max(K, p.K);
        nch = std::max(nch, (int)p.chunks.size());
        work = std::max(work, p.work);
        AdjPlan q;
        if (plan_adjoint(c, c->incl[items[ii].incl], items[ii].nfr, true, q)) tape_ok = 0;
    }
    c->err = keep_err;
    long long nodes = 0;
    for (const RaySet &r : c->incl) nodes = std::max(nodes, r.nnodes);
    out[0] = (double)cuts.size();
    out[1] = (double)most;
    out[2] = (double)K;
    out[3] = (double)nch;
    out[4] = (double)work;
    out[5] = (double)ctx_budget(c);
    out[6] = (double)tape_ok;
    out[7] = (double)nodes;
    return 0;
}

extern "C" int rl_memory_info(rl_ctx *c, double *out) {
    if (!c || !out) return RL_E_ARG;
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    out[0] = (double)ctx_budget(c);
    out[1] = (double)resident_traj(c);
    out[2] = (double)c->last_work;
    out[3] = (double)fr;
    out[4] = (double)tot;
    return 0;
}

// ---------------------------------------------------------------------
// device pre-pass (pyradlite's per-cell tables) and its adjoint
// ---------------------------------------------------------------------
extern "C" int rl_set_level_data(rl_ctx *c, int nlev, const double *E_K, const double *g) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || nlev < 2 || !E_K || !g) return fail(c, RL_E_ARG, "rl_set_level_data: bad args");
    if (c->nlev && nlev != c->nlev)
        return fail(c, 13, "rl_set_level_data: %d levels, the populations have %d", nlev,
                    c->nlev);
    HIPCHK(c, hipSetDevice(c->device));
    if (int e = upload(c, c->d_EK, E_K, (size_t)nlev * 8)) return e;
    if (int e = upload(c, c->d_g, g, (size_t)nlev * 8)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->pre_nlev = nlev;
    return 0;
}

extern "C" int rl_set_dust_species(rl_ctx *c, int nspec, const double *rho_d,
                                   const double *kabs_nu0, const double *kscat_nu0) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !c->nr || !c->nlines || nspec < 0 || (nspec && (!rho_d || !kabs_nu0 || !kscat_nu0)))
        return fail(c, RL_E_ARG, "rl_set_dust_species: set the grid and lines first; bad args");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t ncell = (size_t)c->nr * c->nh;
    if (nspec) {
        if (int e = upload(c, c->d_rhod, rho_d, (size_t)nspec * ncell * 8)) return e;
        if (int e = upload(c, c->d_kabs, kabs_nu0, (size_t)nspec * c->nlines * 8)) return e;
        if (int e = upload(c, c->d_kscat, kscat_nu0, (size_t)nspec * c->nlines * 8)) return e;
    }
    std::vector<double> nu(c->nu0);
    if (int e = upload(c, c->d_nu0, nu.data(), nu.size() * 8)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->pre_nspec = nspec;
    return 0;
}

extern "C" int rl_set_dust_scatter(rl_ctx *c, int nfreq, const double *cont_nu,
                                   const double *scat) {
    if (c) c->gen++;
    if (!c || !c->nr || !c->nlines || nfreq < 0 || (nfreq && (nfreq < 2 || !cont_nu || !scat)))
        return fail(c, RL_E_ARG, "rl_set_dust_scatter: set the grid and lines first; bad args");
    HIPCHK(c, hipSetDevice(c->device));
    c->scat_nf = nfreq;
    if (!nfreq) return 0;
    const size_t ncell = (size_t)c->nr * c->nh;
    // each line centre's continuum interval and weight (the hunt of
    // line.F:3565 on the grid given here)
    const double *x = cont_nu - 1;
    const bool asc = x[nfreq] > x[1];
    std::vector<int> lo_l(c->nlines);
    std::vector<double> w_l(c->nlines, 0.0);
    for (int il = 0; il < c->nlines; il++) {
        const double freq = c->nu0[il];
        int lo = 0, hi = nfreq + 1;
        while (hi - lo != 1) {
            int m = (hi + lo) >> 1;
            if ((freq > x[m]) == asc) lo = m; else hi = m;
        }
        lo_l[il] = lo;
        if (lo != 0 && lo != nfreq) w_l[il] = (freq - x[lo]) / (x[lo + 1] - x[lo]);
    }
    if (int e = upload(c, c->d_scat, scat, (size_t)nfreq * ncell * 8)) return e;
    if (int e = upload(c, c->d_scat_lo, lo_l.data(), lo_l.size() * 4)) return e;
    if (int e = upload(c, c->d_scat_w, w_l.data(), w_l.size() * 8)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static void prepass_args(rl_ctx *c, rl::PrepassLaunch &a) {
    a.ncell = c->nr * c->nh;
    a.nlev = c->pre_nlev;
    a.nlines = c->nlines;
    a.nspec = c->pre_nspec;
    a.dust_from_gas = c->pre_dust_from_gas;
    a.T = c->d_T.as<double>();
    a.la = c->d_la.n ? c->d_la.as<double>() : nullptr;
    a.vt = c->pre_vt_explicit ? c->d_vt.as<double>() : nullptr;
    a.Tw = c->d_Td.as<double>();             // species 0, or the gas temperature's copy
    a.EK = c->d_EK.as<double>();
    a.g = c->d_g.as<double>();
    a.q = c->pre_q;
    a.nq = c->pre_nq;
    a.qx = c->d_qx.as<double>();
    a.qy = c->d_qy.as<double>();
    a.cw = c->pre_cw;
    a.nu0 = c->d_nu0.as<double>();
    a.rho_d = c->d_rhod.as<double>();
    a.Td = c->d_Td.as<double>();
    a.kabs = c->d_kabs.as<double>();
    a.kscat = c->d_kscat.as<double>();
    a.scat = c->scat_nf ? c->d_scat.as<double>() : nullptr;
    a.scat_lo = c->d_scat_lo.as<int>();
    a.scat_w = c->d_scat_w.as<double>();
    a.scat_nf = c->scat_nf;
    a.pops = c->d_pops.as<double>();
    a.lw = c->d_lw.as<double>();
    a.abund = c->d_abund.as<double>();
    a.jd = c->d_jd.as<double>();
    a.ad = c->d_ad.as<double>();
    a.part = c->d_part.as<double>();
    a.g_pops = c->d_gpops.as<double>();
    a.g_lw = c->d_glw.as<double>();
    a.g_abund = c->d_gabund.as<double>();
    a.g_jd = c->d_gjd.as<double>();
    a.gT = c->d_gT.as<double>();
    a.gTd = c->d_gTd.as<double>();
    a.gvt = c->d_gvt.as<double>();
    a.gla = c->d_gla.as<double>();
}

// pyradlite's constants (astropy's CODATA 2018 values, radlite.py:43-51)
constexpr double PYR_KB = 1.380649e-16, PYR_MP = 1.67262192369e-24, PYR_G = 6.6743e-8;

extern "C" int rl_set_partition_table(rl_ctx *c, int n, const double *T, const double *Q) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || n < 0 || n == 1 || (n > 0 && (!T || !Q)))
        return fail(c, RL_E_ARG, "rl_set_partition_table: bad args");
    HIPCHK(c, hipSetDevice(c->device));
    for (int i = 1; i < n; i++)
        if (!(T[i] > T[i - 1]))
            return fail(c, RL_E_ARG, "rl_set_partition_table: temperatures not increasing");
    c->pre_nq = n;
    if (n == 0) {
        c->d_qx = DevBuf();
        c->d_qy = DevBuf();
        return 0;
    }
    if (int e = upload(c, c->d_qx, T, (size_t)n * 8)) return e;
    if (int e = upload(c, c->d_qy, Q, (size_t)n * 8)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_set_turbulence_alpha(rl_ctx *c, double alpha, double gamma, double mu) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c) return RL_E_ARG;
    if (alpha < 0.0) {
        c->pre_alpha_on = 0;
        return 0;
    }
    if (!(gamma > 0.0) || !(mu > 0.0))
        return fail(c, RL_E_ARG, "rl_set_turbulence_alpha: bad args");
    c->pre_alpha_on = 1;
    c->pre_alpha = alpha;
    c->pre_gamma = gamma;
    c->pre_mu = mu;
    return 0;
}

extern "C" int rl_set_velocity_kepler(rl_ctx *c, double mstar) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !(mstar > 0.0)) return fail(c, RL_E_ARG, "rl_set_velocity_kepler: bad args");
    if (!c->have_medium)
        return fail(c, 13, "rl_set_velocity_kepler: set the medium first");
    HIPCHK(c, hipSetDevice(c->device));
    const int ncell = c->nr * c->nh;
    HIPCHK(c, c->d_vel.ensure((size_t)ncell * 24));
    // the radii r_1..r_nr sit at index 2.. of the padded grid (rl_set_grid)
    rl::launch_kepler(c->stream, ncell, c->nh, c->d_rc.as<double>() + 2, PYR_G * mstar,
                      c->d_vel.as<double>());
    HIPCHK(c, hipGetLastError());
    c->h_vel.clear();                        // the next rl_set_medium uploads its own
    for (auto &s : c->incl) s.traj_ready = false;        // u depends on the velocity
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_set_physics(rl_ctx *c, const double *T_gas, const double *T_dust,
                              const double *log_abund, const double *vturb_kms, double m_mol,
                              double q_over) {
    if (c) c->gen++;                         // invalidates a tape
    if (!c || !T_gas || (!vturb_kms && !c->pre_alpha_on) || !(m_mol > 0.0))
        return fail(c, RL_E_ARG, "rl_set_physics: bad args");
    if (!c->have_medium || !c->nlines || !c->pre_nlev || !c->d_nu0.n)
        return fail(c, 13, "rl_set_physics: set the medium, lines, level data and dust species "
                           "first");
    if (q_over < 0.0 && c->pre_nq < 2)
        return fail(c, 13, "rl_set_physics: a tabulated partition sum (q < 0) needs "
                           "rl_set_partition_table");
    if (c->pre_alpha_on && T_dust && c->pre_nspec < 1)
        return fail(c, 13, "rl_set_physics: the alpha model takes the first dust species' "
                           "temperature, and there is no dust species");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t ncell = (size_t)c->nr * c->nh, nsp = (size_t)c->pre_nspec;
    if (int e = upload(c, c->d_T, T_gas, ncell * 8)) return e;
    // explicit turbulence unless the alpha model is on
    c->pre_vt_explicit = c->pre_alpha_on ? 0 : 1;
    if (c->pre_vt_explicit) {
        if (int e = upload(c, c->d_vt, vturb_kms, ncell * 8)) return e;
    } else {
        HIPCHK(c, c->d_vt.ensure(ncell * 8));
    }
    if (log_abund) {
        if (int e = upload(c, c->d_la, log_abund, ncell * 8)) return e;
    } else {
        c->d_la = DevBuf();
    }
    HIPCHK(c, c->d_Td.ensure(std::max(nsp, (size_t)1) * ncell * 8));
    c->pre_dust_from_gas = T_dust ? 0 : 1;
    if (nsp == 0)                            // the width's temperature (alpha model)
        HIPCHK(c, hipMemcpyAsync(c->d_Td.p, c->d_T.p, ncell * 8, hipMemcpyDeviceToDevice,
                                 c->stream));
    for (size_t s = 0; s < nsp; s++) {
        if (T_dust)
            HIPCHK(c, hipMemcpyAsync(c->d_Td.as<double>() + s * ncell, T_dust + s * ncell,
                                     ncell * 8, hipMemcpyHostToDevice, c->stream));
        else
            HIPCHK(c, hipMemcpyAsync(c->d_Td.as<double>() + s * ncell, c->d_T.p, ncell * 8,
                                     hipMemcpyDeviceToDevice, c->stream));
    }
    c->nlev = c->pre_nlev;
    HIPCHK(c, c->d_pops.ensure(ncell * c->nlev * 8));
    HIPCHK(c, c->d_lw.ensure(ncell * 8));
    c->lw_gen++;                             // the device pre-pass rewrites the widths
    HIPCHK(c, c->d_abund.ensure(ncell * 8));
    HIPCHK(c, c->d_jd.ensure(ncell * c->nlines * 8));
    HIPCHK(c, c->d_ad.ensure(ncell * c->nlines * 8));
    const int nblk = (int)((ncell + 255) / 256);
    HIPCHK(c, c->d_part.ensure((size_t)nblk * 8));
    c->pre_q = q_over;
    // thermal broadening 2k/(m_mol m_p) in (km/s)^2 per K (radlite.py:1620); the
    // alpha model adds (alpha c_s)^2 / T = alpha^2 gamma k / (mu m_p) (:1617-1618)
    c->pre_cw = 2.0 * PYR_KB / (m_mol * PYR_MP) / 1e10;
    if (c->pre_alpha_on)
        c->pre_cw += c->pre_alpha * c->pre_alpha * (c->pre_gamma * PYR_KB / (c->pre_mu * PYR_MP)) /
                     1e10;
    rl::PrepassLaunch a;
    prepass_args(c, a);
    rl::launch_prepass(c->stream, a);
    HIPCHK(c, hipGetLastError());
    std::vector<double> part(nblk);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->d_part.p, (size_t)nblk * 8, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double mx = -1.0e300;
    for (double v : part) mx = std::max(mx, v);
    c->aksmax = mx;                          // line.F:3020-3033 (widths > 0 here)
    c->have_dust = true;
    c->pre_ready = 1;
    c->cells_line = -1;
    return 0;
}

extern "C" int rl_physics_adjoint(rl_ctx *c, double *g_T_gas, double *g_T_dust,
                                  double *g_log_abund, double *g_vturb) {
    if (!c) return RL_E_ARG;
    if (!c->pre_ready || !c->d_gpops.n)
        return fail(c, 13, "rl_physics_adjoint: needs rl_set_physics and then "
                           "rl_trace_lines_adjoint");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t ncell = (size_t)c->nr * c->nh, nsp = (size_t)c->pre_nspec;
    HIPCHK(c, c->d_gT.ensure(ncell * 8));
    HIPCHK(c, c->d_gTd.ensure(std::max(nsp, (size_t)1) * ncell * 8));
    HIPCHK(c, c->d_gvt.ensure(ncell * 8));
    HIPCHK(c, c->d_gla.ensure(ncell * 8));
    rl::PrepassLaunch a;
    prepass_args(c, a);
    rl::launch_prepass_adjoint(c->stream, a);
    HIPCHK(c, hipGetLastError());
    auto down = [&](double *dst, DevBuf &b, size_t n) -> int {
        if (dst) HIPCHK(c, hipMemcpyAsync(dst, b.p, n * 8, hipMemcpyDeviceToHost, c->stream));
        return 0;
    };
    if (int e = down(g_T_gas, c->d_gT, ncell)) return e;
    if (g_T_dust && !c->pre_dust_from_gas)
        if (int e = down(g_T_dust, c->d_gTd, nsp * ncell)) return e;
    if (int e = down(g_log_abund, c->d_gla, ncell)) return e;
    if (int e = down(g_vturb, c->d_gvt, ncell)) return e;    // 0 under the alpha model
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int rl_get_tables(rl_ctx *c, double *pops, double *lw, double *abund, double *jd,
                             double *ad) {
    if (!c || !c->nr) return RL_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    const size_t ncell = (size_t)c->nr * c->nh;
    auto down = [&](double *dst, DevBuf &b, size_t n) -> int {
        if (dst) {
            if (b.n < n * 8) return fail(c, 13, "rl_get_tables: table not set");
            HIPCHK(c, hipMemcpyAsync(dst, b.p, n * 8, hipMemcpyDeviceToHost, c->stream));
        }
        return 0;
    };
    if (int e = down(pops, c->d_pops, ncell * c->nlev)) return e;
    if (int e = down(lw, c->d_lw, ncell)) return e;
    if (int e = down(abund, c->d_abund, ncell)) return e;
    if (int e = down(jd, c->d_jd, ncell * c->nlines)) return e;
    if (int e = down(ad, c->d_ad, ncell * c->nlines)) return e;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------------
// host pre-pass (R13) with the reference's arithmetic
// ---------------------------------------------------------------------
extern "C" int rl_prepare_line_freqs(int nl, const double *E, const int *lu, const int *ll,
                                     double *nu0) {
    if (nl < 1 || !E || !lu || !ll || !nu0) return RL_E_ARG;
    for (int i = 0; i < nl; i++) {                      // line.F:1903, 1980-1983
        double eu = 1.986468498e-16 * E[lu[i]], el = 1.986468498e-16 * E[ll[i]];
        nu0[i] = 1.509160e26 * (eu - el);
    }
    return 0;
}

static double bplanck(double temp, double nu) {          // setup.F:937-952
    if (temp == 0.0) return 0.0;
    return 1.47455e-47 * nu * nu * nu / (std::exp(4.7989e-11 * nu / temp) - 1.0) + 1.0e-290;
}

extern "C" int rl_prepare_line_dust_scat(int nr, int nh, int nspec, const double *rho_d,
                                         const double *t_d, int nf, const double *cn,
                                         const double *kabs, const double *kscat,
                                         const double *scat, int nl, const double *nu0,
                                         double *jd, double *ad) {
    if (nr < 1 || nh < 1 || nspec < 0 || nf < 1 || nl < 1 || !cn || !nu0 || !jd || !ad)
        return RL_E_ARG;
    // line_dust_compute_src_alp, line.F:3560-3607: the continuum interval of
    // each line centre (hunt), its weight and the interpolated opacities are
    // the same for every cell, so they are formed once per line with the
    // reference's expressions; cells are split over host threads
    const double *x = cn - 1;
    const bool asc = x[nf] > x[1];
    const size_t ncell = (size_t)nr * nh;
    std::vector<int> lo_l(nl);
    std::vector<double> wgt_l(nl, 0.0), ka_l((size_t)nl * nspec), ks_l((size_t)nl * nspec);
    for (int il = 0; il < nl; il++) {
        const double freq = nu0[il];
        int lo = 0, hi = nf + 1;
        while (hi - lo != 1) {
            int m = (hi + lo) >> 1;
            if ((freq > x[m]) == asc) lo = m; else hi = m;
        }
        lo_l[il] = lo;
        if (lo == 0 || lo == nf) continue;
        const double wgt = (freq - x[lo]) / (x[lo + 1] - x[lo]);
        wgt_l[il] = wgt;
        for (int sp = 0; sp < nspec; sp++) {
            const double *ka = kabs + (size_t)sp * nf - 1, *ks = kscat + (size_t)sp * nf - 1;
            ka_l[(size_t)il * nspec + sp] = wgt * ka[lo + 1] + (1.0 - wgt) * ka[lo];
            ks_l[(size_t)il * nspec + sp] = wgt * ks[lo + 1] + (1.0 - wgt) * ks[lo];
        }
    }
    auto work = [&](size_t c0, size_t c1) {
        for (size_t cell = c0; cell < c1; cell++) {
            for (int il = 0; il < nl; il++) {
                const int lo = lo_l[il];
                double src = 0.0, alp = 0.0;
                if (lo != 0 && lo != nf) {
                    const double freq = nu0[il], wgt = wgt_l[il];
                    for (int sp = 0; sp < nspec; sp++) {
                        const double kw = ka_l[(size_t)il * nspec + sp];
                        const double rd = rho_d[sp * ncell + cell];
                        src = src + rd * kw * bplanck(t_d[sp * ncell + cell], freq);
                        alp = alp + rd * kw;
                    }
                    // scattering/PAH source (scatsource.dat + quantsource.dat,
                    // source.F:494-546), line.F:3593-3594; 0 without the files
                    const double s1 = scat ? scat[(size_t)lo * ncell + cell] : 0.0;
                    const double s0 = scat ? scat[(size_t)(lo - 1) * ncell + cell] : 0.0;
                    src = src + wgt * s1 + (1.0 - wgt) * s0;
                    for (int sp = 0; sp < nspec; sp++)
                        alp = alp + rho_d[sp * ncell + cell] * ks_l[(size_t)il * nspec + sp];
                }
                jd[cell * nl + il] = src;
                ad[cell * nl + il] = alp;
            }
        }
    };
    const size_t work_total = ncell * (size_t)nl;
    unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (work_total < ((size_t)1 << 16)) nt = 1;
    std::vector<std::thread> th;
    const size_t per = (ncell + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const size_t c0 = t * per, c1 = std::min(ncell, c0 + per);
        if (c0 < c1) th.emplace_back(work, c0, c1);
    }
    for (auto &t : th) t.join();
    return 0;
}

extern "C" int rl_prepare_line_dust(int nr, int nh, int nspec, const double *rho_d,
                                    const double *t_d, int nf, const double *cn,
                                    const double *kabs, const double *kscat, int nl,
                                    const double *nu0, double *jd, double *ad) {
    return rl_prepare_line_dust_scat(nr, nh, nspec, rho_d, t_d, nf, cn, kabs, kscat, nullptr, nl,
                                     nu0, jd, ad);
}

extern "C" int rl_prepare_star(int nf, const double *fnu, double rstar, double *is) {
    if (nf < 0 || (nf && (!fnu || !is)) || !(rstar > 0.0)) return RL_E_ARG;
    for (int i = 0; i < nf; i++) is[i] = 3.0308410e36 * fnu[i] / (rstar * rstar);
    return 0;
}

extern "C" double rl_prepare_aksmax(int n, const double *lw, double fix_lw) {
    double mx = 0.0;                                     // line.F:3020-3033
    for (int i = 0; i < n; i++) {
        double a = lw[i];
        if (a <= 0.0) a = fix_lw;
        if (a > mx) mx = a;
    }
    return mx;
}

extern "C" int rl_passband(double vmax, double dv, int *nfr, double *vmax_out) {
    if (!(dv > 0.0) || !(vmax > 0.0) || !nfr || !vmax_out) return RL_E_ARG;
    int n = (int)(2.0 * vmax / dv + 1);                  // main.F:209-210
    *nfr = n;
    *vmax_out = 0.5 * (n - 1) * dv;
    return n >= 2 ? 0 : 13;
}
