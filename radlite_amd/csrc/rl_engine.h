// Host/device interface between the C ABI (rl_api.cpp) and the kernels
// (rl_engine.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "rl_fwd32.h"

namespace rl {

struct GridDev {
    int nr, ny;
    const double *rc, *tc;    // ghost-extended, index i+1
    double rstar;
    int t_sorted;             // tc non-decreasing over 1..ny: the node pass may take its
                              // theta cell from the crossing pass's guess
};

// Channel-independent geometry of one trajectory node (FP64), 48 bytes.
struct GeoNode {
    double ds;                // s_k - s_{k-1} (0 for the first node)
    double u;                 // Omega.v/c at the node (line.F:2650-2662)
    double dr, dt;            // interpolation fractions (telescope.F:4088-4091)
    int c00, c01, c10;        // stencil cells (r0,t0) (r0,t1) (r1,t0)
    int flag;                 // icross | node flags << 2
    // the fourth cell (r1,t1): cell(ir, it) = (ir-1) nh + (it-1) is linear,
    // so cell(r1,t1) = cell(r1,t0) + cell(r0,t1) - cell(r0,t0)
    __host__ __device__ int c11() const { return c10 + c01 - c00; }
};
static_assert(sizeof(GeoNode) == 48, "GeoNode layout");

enum { CR_JD = 0, CR_AD = 1, CR_LW = 2, CR_NU = 3, CR_NL = 4 };
// velocity and line width of a cell, and a ray's start position, packed
// for the FP32 node pass
struct alignas(16) VelLw { double vx, vy, vz, lw; };
struct alignas(16) RayXZ { double x0, z0; };
struct CellRec { double v[5]; };

struct RayInfo {
    float mn, mx;             // NONREDUNDANT window (REAL*4)
    int cont;                 // first skipped channel, -1 if none
    int n;                    // nodes on the ray
};

struct LineConst {
    double nu0, c_aa, hknu, A, Bud, Bdu;
    double starfract;         // (R*/rbeam0)^2 for the central beam
};

// One item of a batched forward trace (device array, one per item).
struct ItemDev {
    int nrays, nfr, nrings, nphi, skip_mode;
    int cslot;                 // cell table slot of the item's line in the batch
    const long long *off;      // node CSR of the item's inclination
    const GeoNode *geo;
    long long hot_off;         // FP32: the inclination's segment records, in bytes from the
    long long aux_off;         // batch's record buffer (so that the kernels see global memory)
    long long rec0;            // first node record
    long long img0;            // first image value ([ray][channel])
    long long chan0;           // per channel: dnu at chan0, velo at chan0+nfr, the star's
                               // intensity at chan0+2nfr, the outer boundary's at chan0+3nfr
    long long part0;           // partial flux [ring][channel]
    long long flux0;
    int rinfo0;
    LineConst lc;
    double aks2;
    const double *ring_w;
    double center_w;
};

struct BatchLaunch {
    int nitems, max_rays, max_nfr, max_rings, ncell, any_skip;
    const ItemDev *items;      // device
    const CellRec *cells;      // [slots][ncell]
    void *rec;
    RayInfo *rinfo;
    const double *chan;
    double *image, *partial, *flux;
    unsigned char *cmask;      // optional [image layout]: 1 where integrated
    double d2;
    const long long *tile0;    // device [nitems+1] first tile of each item
    long long ntiles;
    unsigned long long *ovs, *counter;
    int *maser;
};

void launch_batch_nodes(hipStream_t st, int fp64, const BatchLaunch &b);
// FP32: the line-independent segment records of a ray set (rl_fwd32.h)
void launch_segrec(hipStream_t st, long long nnodes, const long long *off, int nrays,
                   const GeoNode *geo, const double *lw, SegHot *hot, SegAux *aux);

// Fixed device workspace (rl_scratch.hip) of the allocation-free forward
// entry point: SCRATCH_BYTES per GPU, resident while the process runs.
// scratch_attach returns its device address on `device` (which must be
// current), looked up once per device.
constexpr size_t SCRATCH_BYTES = (size_t)256 << 20;
constexpr int SCRATCH_MAX_DEVICES = 64;
hipError_t scratch_attach(int device, void **out);
// nblocks == 0: static grid over the first ntiles tiles; > 0: persistent
// grid of nblocks blocks taking tiles from the work counter
void launch_batch_formal(hipStream_t st, int fp64, const BatchLaunch &b, int cw, int nblocks,
                         long long ntiles);
void launch_batch_finish(hipStream_t st, int fp64, const BatchLaunch &b);
void launch_batch_flux(hipStream_t st, const BatchLaunch &b);
void launch_zero_u64(hipStream_t st, unsigned long long *p, int n);
void launch_rays_circular(hipStream_t st, int nrings, int nphi, const double *tab, double *x0,
                          double *z0);
int formal_batch_blocks_per_cu(int fp64);

// One item of the adjoint (scaled records of precision T, all channels).
// A chunk is a range of the ray set's rays: nrays, nnodes and every
// per-node and per-ray buffer are the chunk's own; off points at the
// chunk's slice of the ray set's node CSR (global node indices), geo at the
// ray set's nodes; ray0 and node0 are the chunk's first ray and node.
struct AdjLaunch {
    int nrays, nfr, nrings, nphi;
    long long nnodes;
    int ray0;
    long long node0;
    int ncell, nlev, nlines, line, lev_up, lev_low;
    double molpg;
    const long long *off, *ckoff;      // node CSR; checkpoint CSR (blocks per ray)
    const GeoNode *geo;
    const CellRec *cells;
    void *rec;                         // FastRecT<T>[nnodes]
    RayInfo *rinfo;
    void *ckpt;                        // T[sum blocks][nfr]
    int kblk;                          // checkpoint interval in segments (1..16)
    const double *dnu, *velo, *istar;  // per channel
    const double *i0;                  // per channel: outer boundary intensity
    LineConst lc;
    double aks2;
    double *image;                     // [nrays][nfr] forward intensities (the chunk's rows;
                                       // the whole item's for launch_adjoint_flux)
    const double *ring_w;
    double center_w;
    int nchunks;
    double *partial, *flux;
    double d2;
    const double *gflux;               // device [nfr] dL/dflux
    double *gnode;                     // [nnodes][5] scaled-unit node gradients
    double *gtrue;                     // [nnodes][5] (jd, ad, lw, N_up, N_low)
    const int *st_off, *st_vals;       // stencil transpose: cell -> entries 4k+j
    const double *st_w;                // entry weights
    const double *rho, *abund, *pops;
    double *g_rho, *g_abund, *g_lw, *g_pops, *g_jd, *g_ad;   // accumulated over items
};

// One ray set's crossing pass: rays, direction, and where its crossings go
// (per-ray slots of nmax entries) and its scratch.
// a crossing of the crossing pass in its ray's slot: s, the packed code
// (icross | rext << 2 | node flags << 3 | index << 5) and the cell guesses
// for the node pass: the indices of the last sphere and the last cone the
// ray crossed before this crossing, each plus one, in the low and high 16
// bits (0: none); 16 bytes, so that 8 records fill a 128-byte line
struct CrossRec {
    double s;
    int code, guess;
};

struct CrossSet {
    int nrays, central_blend;
    const double *x0, *z0;
    const double *bpix;        // rectangular images: pixel radius per ray, or null
    double bmax;               // rays with bpix >= bmax are not traced (0 nodes)
    double theta0, costh0, sinth0;
    double *s1, *s2;
    int *iy;
    int *count;
    CrossRec *cr;              // [ray][nmax]
};
constexpr int CROSS_SETS_MAX = 8;
struct CrossSets {
    CrossSet s[CROSS_SETS_MAX];
    int n;
};

void launch_cross(hipStream_t st, const GridDev &g, const double *tan2, const CrossSets &sets,
                  int nmax, int *err);
// NF_STAR on the nodes of the listed rays (impact parameter within R*)
// that follow a node of radial cell 1 and are in it themselves
void launch_star_flags(hipStream_t st, int nlisted, const int *rays, const long long *off,
                       GeoNode *nodes, int nh);

// Position-velocity cube of one item: every (ray, channel) integrated,
// intensity and line-of-sight optical depth; rays without nodes (outside
// the grid) give 0 and 0.
struct PosvelLaunch {
    int nrays, nfr, nx, ny;            // rays: nx*ny pixels (ix-major), then extra rays
    const long long *off;
    const GeoNode *geo;
    const CellRec *cells;
    RayInfo *rinfo;
    double aks2;
    void *rec;                         // NodeRec<double> (FP64) or FastRec (FP32)
    const double *dnu, *velo, *istar;  // per channel
    const double *i0;                  // per channel: outer boundary intensity
    LineConst lc;
    double *image, *tau;               // [iy][ix][channel]
    double *extra;                     // [ray - nx*ny][channel] (intensity)
    unsigned long long *ovs;
    int *maser;
};
template <typename T> void launch_posvel(hipStream_t st, const PosvelLaunch &a);
void launch_geo(hipStream_t st, const GridDev &g, const double *vel, int nh, int nrays,
                int nmax, const long long *off, const double *x0, const double *z0,
                double costh0, double sinth0, const CrossRec *cr,
                GeoNode *nodes, int *err, long long nnodes = 0, const double *lw = nullptr,
                SegHot *hot = nullptr, SegAux *aux = nullptr,
                VelLw *vlw = nullptr,     // hot != null: the FP32 segment records too, the
                RayXZ *xz = nullptr);     // velocities and widths packed into vlw [cell],
                                          // the rays' x0, z0 into xz [ray]
void launch_cells(hipStream_t st, int ncell, int nlev, int nlines, int line, int lev_up,
                  int lev_low, double molpg, const double *rho, const double *abund,
                  const double *lw, const double *pops, const double *jd, const double *ad,
                  CellRec *out);
template <typename T> void launch_adjoint_fwd(hipStream_t st, const AdjLaunch &a);
template <typename T> void launch_adjoint_rev(hipStream_t st, const AdjLaunch &a);
void launch_adjoint_flux(hipStream_t st, const AdjLaunch &a);
int adjoint_block_len();
// Stencil transpose (cell -> node entries) of a node set; with temp ==
// nullptr only *temp_bytes is set.  keys/vals: 4*nnodes ints each (a, b);
// w: 4*nnodes doubles; off: ncell+1 ints.
int build_stencil_transpose(hipStream_t st, long long nnodes, const GeoNode *geo, int ncell,
                            int *keys_a, int *keys_b, int *vals_a, int *vals_b, double *w,
                            int *off, void *temp, size_t *temp_bytes);
size_t fast_rec_size(int precision_fp64);

// The device pre-pass (rl_prepass.hip): cell tables from per-cell
// temperature, log-abundance and turbulence, and its adjoint.
struct PrepassLaunch {
    int ncell, nlev, nlines, nspec, dust_from_gas;
    const double *T, *la, *vt;         // [cell]; la may be null (abundance kept); vt null:
    const double *Tw;                  // the alpha model, whose width follows Tw [cell]
    const double *EK, *g;              // level energies [K] and weights [nlev]
    double q;                          // Q = T/q (q > 0), renormalised (q = 0), table (q < 0)
    int nq;                            // partition sum table: nq temperatures qx (ascending)
    const double *qx, *qy;             // and sums qy
    double cw;                         // lw^2 per kelvin: 2k/(m_mol m_p)/1e10 (+ the alpha
                                       // model's (alpha c_s)^2 / T_d / 1e10)
    const double *nu0;                 // [line]
    const double *rho_d, *Td;          // [species][cell]
    const double *kabs, *kscat;        // at the line centres [species][line]
    const double *scat;                // scattering source [continuum freq][cell], or null
    const int *scat_lo;                // per line: its continuum interval (1-based; 0 or
    const double *scat_w;              // nfreq: none) and the weight of its upper end
    int scat_nf;
    double *pops, *lw, *abund, *jd, *ad;
    double *part;                      // block maxima of lw, (ncell+255)/256
    // adjoint
    const double *g_pops, *g_lw, *g_abund, *g_jd;
    double *gT, *gTd, *gvt, *gla;
};
void launch_prepass(hipStream_t st, const PrepassLaunch &a);
// Keplerian velocity [cell][3] about a mass with G M = gm; r = the radii [nr]
void launch_kepler(hipStream_t st, int ncell, int nh, const double *r, double gm, double *vel);
void launch_prepass_adjoint(hipStream_t st, const PrepassLaunch &a);

// Resampling of a circular image onto a grid (rl_resample.hip): out[ch][p]
// = sum_k w[p][k] img[vtx[p][k]][ch] over nv vertices per pixel (vtx < 0:
// no contribution)
void launch_rect_gather(hipStream_t st, int npix, int nfr, int nv, const int *vtx,
                        const double *w, const double *img, double *out);

}  // namespace rl
