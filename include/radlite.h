/*
 * radlite.h -- C ABI of the MI355X line raytracer.
 *
 * Drop-in for RADLite's image/spectrum raytrace: the per-line driver
 * calc_line_spectrum (RADLITE/telescope.F:1465) and everything below it
 * (make_image_circular :339, make_trajectory_c :2787, charintline
 * :3889, get_line_dust_values line.F:3965, clever_integrate_element_
 * linedust line.F:4636, integrate_element_linedust line.F:4515,
 * voigt_profile line.F:2280, qdr_src_2 transfer.F:1498, and the flux
 * sum calc_freq_flux_observer telescope.F:1320).
 *
 * The reference passes everything through COMMON blocks; here the same
 * quantities are passed explicitly, as plain pointers and sizes.  All
 * host arrays are caller-owned and only read during the call (they are
 * copied to device memory).  Cell fields are [nr][nth_half] with theta
 * fastest -- the order RADLite's readers fill them (setup.F:778-785).
 *
 * Return values: 0 on success, otherwise the reference's stop code for
 * the same condition (13 input/capacity, 124 star resolved by the
 * central beam, 749 ds<0, 7454/137 hole logic, 6023/6024 interpolation
 * fraction out of [0,1], 192/987/988 trajectory extra points), or
 * RL_E_HIP for a device-runtime failure.  rl_last_error() describes the
 * last failure.  One host thread per context; contexts are independent.
 */
#ifndef RADLITE_H
#define RADLITE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rl_ctx rl_ctx;

enum { RL_FP64 = 0, RL_FP32 = 1 };
/* RL_SKIP_REFERENCE reproduces the reference's NONREDUNDANT channel
 * skipping (telescope.F:550-613); RL_SKIP_NONE integrates every channel. */
enum { RL_SKIP_REFERENCE = 0, RL_SKIP_NONE = 1 };
enum { RL_WARN_MASER = 1u, RL_WARN_OVERSHOOT = 2u };
enum { RL_E_HIP = 900, RL_E_ARG = 13 };

/* One (line, inclination) work item.  line: 0-based line index;
 * incl: 0-based index into the inclinations given to
 * rl_setup_rays_circular; nfr: channels (>=2); vmax_kms: half width of
 * the passband (telescope.F:1465 argument "passband"). */
typedef struct {
    int line;
    int incl;
    int nfr;
    double vmax_kms;
    int skip_mode;
} rl_item;

/* Context: owns device buffers and a stream on `device`.  precision is
 * RL_FP64 (parity mode) or RL_FP32 (performance mode; FP64 is kept for
 * the trajectory roots and path lengths). */
int  rl_create(rl_ctx **ctx, int device, int precision);
void rl_destroy(rl_ctx *ctx);
const char *rl_last_error(const rl_ctx *ctx);

/* Spatial grid: radius.inp radii [nr] and theta.inp values [nth_half];
 * mirrored and given ghost cells as grid.F:711-819,1098-1176 do. */
int rl_set_grid(rl_ctx *ctx, int nr, const double *r_cm, int nth_half,
                const double *theta_half);

/* Medium (setup.F:754-911,1453): gas density [g/cm^3], molecular
 * abundance, velocity [nr][nth][3] (v_r, v_theta, v_phi in cm/s), local
 * line width lw [km/s] (turbulence.inp), mean molecular weight, and the
 * maximum line width aksmax [km/s] (line.F:3020-3033). */
int rl_set_medium(rl_ctx *ctx, const double *rho, const double *abund,
                  const double *vel_cms, const double *lw_kms,
                  double umass_av, double aksmax_kms);

/* Level populations [nr][nth][nlev] (line.F:915-919). */
int rl_set_levels(rl_ctx *ctx, int nlev, const double *pops);

/* Lines: rest frequency nu0 [Hz] (line.F:1980-1983), Einstein A, upper
 * and lower statistical weights, 0-based level indices. */
int rl_set_lines(rl_ctx *ctx, int nlines, const double *nu0,
                 const double *A, const double *gu, const double *gl,
                 const int *lev_up, const int *lev_low);

/* Grey dust emissivity/extinction at each line centre,
 * [nr][nth][nlines] (line.F:3687-3743). */
int rl_set_dust_line(rl_ctx *ctx, const double *jd, const double *ad);

/* Star: radius and specific intensity at the stellar surface on the
 * continuum grid cont_nu (star.F:508-525); ncont may be 0 (dark star). */
int rl_set_star(rl_ctx *ctx, double rstar_cm, int ncont,
                const double *cont_nu, const double *istar);

/* Outer boundary (radlite.inp, charintline telescope.F:3996-4030):
 * 0 = black, 2 = CMB at the line centre, 3 = the interstellar radiation
 * field of rl_set_interstellar_field at each channel frequency; 1 and
 * other values are the reference's stop 13. */
int rl_set_outer_bc(rl_ctx *ctx, int type);

/* Interstellar radiation field I_nu on the continuum grid
 * (interstellfield.inp, read_interstellfield star.F:675-720), for outer
 * boundary type 3: interpolated to every channel of a circular-image
 * spectrum as line_init_passband_outerbc (line.F:3855-3902) does,
 * including its test of the continuum interval against the channel count
 * (a channel in interval nfr starts from 0).  Position-velocity cubes
 * start from 0 with type 3, as calc_write_line_posvel leaves the field
 * unset.  A trace with type 3 and no field is stop 13. */
int rl_set_interstellar_field(rl_ctx *ctx, int ncont, const double *cont_nu,
                              const double *isrf);

/* Circular-image ray sets, one per inclination (setup_rays_circular,
 * telescope.F:715-1191 with imethod=1, nrref=10).  Outputs (optional,
 * may be NULL): per inclination the ray and ring counts, and the ring
 * radii [nrings+1] and ring edges [nrings+2] of the first inclination. */
int rl_setup_rays_circular(rl_ctx *ctx, int nincl, const double *incl_rad,
                           int nphi, int nrext, int dbdr, int *n_rays,
                           int *n_rings, double *ring_r, double *ring_edge);

/* Trace a batch of (line, inclination) items.  flux: concatenation of
 * every item's nfr values in erg/s/cm^2/Hz at 1 pc (telescope.F:1435).
 * image (optional): concatenation per item of [nrays][nfr] intensities
 * (ray 0 is the central beam, then ring-major, phi fastest).
 * warn_bits (optional): per item RL_WARN_* flags. */
int rl_trace_lines_forward(rl_ctx *ctx, int nitems, const rl_item *items,
                           double *flux, double *image, unsigned *warn_bits);

/* rl_trace_lines_forward with the integrated-cell mask: cmask (optional,
 * image layout) receives 1 where a (ray, channel) was integrated and 0
 * where NONREDUNDANT skipping copied the continuum (imcir_cmask,
 * telescope.F:550-613, for this call's items only; the reference keeps
 * the mask of earlier lines, which a caller reproduces by OR-ing). */
int rl_trace_lines_forward_ex(rl_ctx *ctx, int nitems, const rl_item *items,
                              double *flux, double *image, unsigned char *cmask,
                              unsigned *warn_bits);

/* rl_trace_lines_forward that allocates no device memory.  No counterpart
 * in the reference: it is for callers that cannot allocate (serving
 * loops, harnesses that forbid hipMalloc).  Every work buffer comes from
 * one fixed workspace of 256 MiB per GPU that the library holds for the
 * life of the process; items that need more (node records, 64 B per
 * trajectory node and item, are most of it) are refused with 13.  flux_dev
 * (device memory, the layout of rl_trace_lines_forward's flux) receives
 * the spectra; image_dev (device memory, optional) the [nrays][nfr]
 * images; warn_bits (host, optional) per item RL_WARN_* flags.  The
 * trajectories of the items' inclinations must be resident, i.e. built by
 * an earlier rl_trace_lines_forward on this context with the same rays
 * and velocity field (13 otherwise).  The items run as one batch; the call
 * returns when its results are in place.  All contexts on a GPU share its
 * workspace: calls that overlap on one GPU race. */
int rl_trace_lines_forward_noalloc(rl_ctx *ctx, int nitems, const rl_item *items,
                                   double *flux_dev, double *image_dev, unsigned *warn_bits);

/* Rectangular-image ray sets for position-velocity cubes, one per
 * inclination (setup_rays_rectang, telescope.F:2229-2468; linespec_command
 * 2 reads nrx, nry, the image size in cm, phioff, xoff, yoff and addstar,
 * telescope.F:195-256).  nx, ny even; an inclination below 0.1 rad is
 * taken as 0.1 (telescope.F:2294-2303); addstar > 0 adds the star when it
 * is smaller than a pixel (then the image must be centred on it).
 * n_rays (optional): per inclination the rays (nx*ny, +1 with addstar).
 * Replaces any circular ray sets. */
int rl_setup_rays_rectang(rl_ctx *ctx, int nincl, const double *incl_rad, int nx,
                          int ny, double size_x_cm, double size_y_cm, double phioff,
                          double xoff_cm, double yoff_cm, int addstar, int *n_rays);

/* Position-velocity cubes of (line, inclination) items on rectangular
 * ray sets (calc_write_line_posvel, telescope.F:1828-2031, with
 * make_freq_image_rectang, telescope.F:2061-2215): image and tau
 * (optional) receive per item [ny][nx][nfr] intensities
 * [erg/s/cm^2/Hz/ster] and optical depths; every channel is integrated
 * (skip_mode is ignored); pixels outside 0.999 of the grid's outer radius
 * are 0 (tau 0); an unresolved star is added to the four central pixels.
 * The reference sets no stellar spectrum on this path, so rays that end
 * on the star start from zero intensity there. */
int rl_trace_lines_posvel(rl_ctx *ctx, int nitems, const rl_item *items,
                          double *image, double *tau, unsigned *warn_bits);

/* Resampling of a circular image onto a rectangular grid: the
 * interpolation step of imcir_conv (PRO/imcir_conv.pro:161-181, IDL
 * triangulate + trigrid).  image: one item's [nrays][nfr] intensities as
 * rl_trace_lines_forward returns them; per output pixel nv vertex rays
 * vtx[npix][nv] (-1: none) and weights w[npix][nv] (the barycentric
 * weights of the pixel in its triangle, times any per-vertex factor); out
 * receives [nfr][npix].  The reference has no such library routine: this
 * replaces the IDL post-processing of the lineposvelcirc file. */
int rl_resample_image(rl_ctx *ctx, int nrays, int nfr, const double *image, int npix,
                      int nv, const int *vtx, const double *w, double *out);

/* Adjoint of the trace for the same items (every item must use
 * RL_SKIP_NONE).  Given dL/dflux (same layout as flux), returns dL/d(the
 * inputs the trace reads) and, in flux (optional), the spectrum.  When
 * the last rl_trace_lines_forward_tape call was on the same items and no
 * input has been set since, only the reverse sweep runs, from that tape;
 * otherwise the checkpointed forward runs first, item by item.  Gradients
 * are summed over the items: g_rho, g_abund, g_lw [nr][nth];
 * g_pops [nr][nth][nlev]; g_jd, g_ad [nr][nth][nlines].  Any output may
 * be NULL; outputs are overwritten.  The reference has no adjoint: this
 * differentiates its formal solution (charintline telescope.F:3889,
 * clever_integrate_element_linedust line.F:4636, qdr_src_2
 * transfer.F:1498) including the sub-grid point positions. */
int rl_trace_lines_adjoint(rl_ctx *ctx, int nitems, const rl_item *items,
                           const double *dL_dflux, double *flux, double *g_rho,
                           double *g_abund, double *g_lw, double *g_pops,
                           double *g_jd, double *g_ad);

/* Forward trace of RL_SKIP_NONE items that keeps, per item, the node
 * records and I before every K-th segment of every (ray, channel) on the
 * device (the tape) for a following rl_trace_lines_adjoint: the forward
 * half of a training step.  flux as in rl_trace_lines_forward.  The tape
 * replaces the previous one and is dropped by rl_tape_release or
 * rl_destroy; any rl_set_* or rl_setup_rays_circular call makes it stale.
 * No counterpart in the reference (its trace has no adjoint). */
int rl_trace_lines_forward_tape(rl_ctx *ctx, int nitems, const rl_item *items,
                                double *flux);
int rl_tape_release(rl_ctx *ctx);

/* Checkpoint interval K (segments) of the last adjoint or tape call,
 * whether the last adjoint call ran from the tape (1) or ran its own
 * forward (0), and the most ray chunks an item of it was cut into (an
 * item whose buffers do not fit the memory budget is traced and reversed
 * a range of rays at a time; the tape needs every item whole). */
int rl_last_adjoint_info(const rl_ctx *ctx, int *ckpt_interval, int *used_tape,
                         int *chunks);

/* Device memory: out[0] the budget, out[1] bytes held by resident
 * trajectories (and adjoint transposes), out[2] work bytes of the last
 * call's largest batch or adjoint chunk (with its image), out[3] free and
 * out[4] total device bytes now.  The budget covers trajectories and a
 * call's work buffers; the trajectories of inclinations a batch does not
 * use are evicted (and rebuilt when needed) to keep within it. */
int rl_memory_info(rl_ctx *ctx, double *out);

/* The plan of a call on these items, without running it (the node counts
 * of their inclinations are computed, their nodes are not built):
 * out[0] forward batches, out[1] bytes of the largest batch with the
 * nodes of its inclinations, out[2] the adjoint's largest checkpoint
 * interval, out[3] the most ray chunks of an adjoint item, out[4] the
 * largest adjoint chunk's bytes with its image, out[5] the budget,
 * out[6] 1 if the adjoint tape of each item alone fits, out[7] the most
 * trajectory nodes of an inclination. */
int rl_plan(rl_ctx *ctx, int nitems, const rl_item *items, double *out);

/* ---- batching controls (no counterpart in the reference) ---------- */
/* A call to rl_trace_lines_forward runs its items in batches: every
 * item of a batch is resident in device memory at once, and each stage
 * (cell tables, node records, formal solution, flux) is one launch for
 * the whole batch, the formal solution as a persistent kernel over a
 * device work counter.  Batches are cut so that their device memory,
 * with the trajectories of their inclinations, stays within the budget
 * (a batch holds at least one item): bytes > 0 sets it, 0 (the default)
 * uses 60% of the device memory free at the context's first trace. */
int rl_set_memory_budget(rl_ctx *ctx, double bytes);

enum {
    RL_OPT_GRAPH = 1,     /* 1 (default): record each batch shape's launches in a
                             hipGraph and replay it; 0: launch directly */
    RL_OPT_AUTOTUNE = 2,  /* 1 (default): time the tile mappings and the grid
                             (static, or persistent over a work counter) once per
                             batch shape and keep the fastest; 0: defaults */
    RL_OPT_TILE = 3,      /* fix the tile mapping (turns the autotuner off):
                             0 = ray-major (ray, channel) pairs, 16/32/64 =
                             channels per ray in a wave */
    RL_OPT_PERSISTENT = 4,/* with a fixed mapping: 1 (default) persistent grid
                             over the work counter, 0 static grid */
    RL_OPT_CKPT = 5       /* adjoint checkpoint interval: 0 (default) chosen
                             from the memory budget (1 when every segment's I
                             fits half of it, else up to 16), or fixed to
                             1, 2, 4, 8 or 16 segments */
};
int rl_set_option(rl_ctx *ctx, int option, int value);

/* Batches of the last forward call, and the tile mapping and grid the
 * last batch used (blocks = 0: static grid; > 0: persistent blocks). */
int rl_last_batch_info(const rl_ctx *ctx, int *batches, int *tile_cw, int *blocks);

/* Wait for all work of the context. */
int rl_sync(rl_ctx *ctx);

/* Milliseconds of device time of the last trace call, by stage:
 * t[0] trajectories, t[1] node records, t[2] formal solution,
 * t[3] flux sum, t[4] total; and counts c[0] rays, c[1] nodes,
 * c[2] integrated ray*channels, c[3] nominal ray*channels. */
int rl_last_stats(const rl_ctx *ctx, double *t_ms, long long *counts);

/* ---- device pre-pass (pyradlite's per-cell tables) ---------------- */
/* The tables the trace reads, computed on the device from per-cell
 * physical fields, as pyradlite's pre-pass forms them before a run
 * (_prep_mol_forcore radlite.py:1024-1127, _calc_turbulence
 * radlite.py:1592-1631, _calc_abundance radlite.py:1481-1531) and
 * RADLite adds the dust (line_dust_compute_src_alp line.F:3560-3607,
 * bplanck setup.F:937-952).  No counterpart in the reference as an entry
 * point: it replaces the host arrays of rl_set_levels, the width and
 * abundance of rl_set_medium and rl_set_dust_line for callers that fit
 * the model, and its adjoint carries the trace adjoint's cell gradients
 * back to the physical fields without leaving the device. */
/* Level energies E_K [K] and statistical weights g [nlev]. */
int rl_set_level_data(rl_ctx *ctx, int nlev, const double *E_K, const double *g);
/* Dust species: density rho_d [nspec][nr][nth] and kabs, kscat at each
 * line centre [nspec][nlines] (the opacity tables interpolated at nu0 as
 * line.F:3560-3607 does; 0 outside the table).  Call after rl_set_lines. */
int rl_set_dust_species(rl_ctx *ctx, int nspec, const double *rho_d,
                        const double *kabs_nu0, const double *kscat_nu0);
/* The dust scattering/PAH source [nfreq][nr][nth] on the continuum grid
 * cont_nu (scatsource.dat + quantsource.dat, source.F:494-546), added by
 * rl_set_physics to j_d at each line centre (line.F:3593-3594); nfreq = 0
 * removes it.  Call after rl_set_lines. */
int rl_set_dust_scatter(rl_ctx *ctx, int nfreq, const double *cont_nu, const double *scat);
/* The partition sum as a table (ParSum.dat's column of the molecule, read
 * by _read_psum radlite.py:1354-1394): n >= 2 temperatures T [K],
 * strictly increasing, and sums Q.  rl_set_physics with q < 0
 * (RL_Q_TABLE) interpolates it linearly at each cell's gas temperature and
 * extrapolates past its ends, as interp1d(kind="linear",
 * fill_value="extrapolate") does at radlite.py:1150-1153.  n = 0 removes
 * it. */
#define RL_Q_TABLE (-1.0)
int rl_set_partition_table(rl_ctx *ctx, int n, const double *T, const double *Q);
/* pyradlite's turbulence model (_calc_turbulence radlite.py:1592-1631):
 * with alpha >= 0, rl_set_physics ignores v_turb (which may be NULL) and
 * sets lw = sqrt((alpha c_s)^2 + 2 k T_d / (m_mol m_p)) with c_s^2 = gamma k
 * T_d / (mu m_p) and T_d the first dust species' temperature (T_gas when
 * rl_set_physics has no T_dust); rl_physics_adjoint then carries the
 * width's gradient to that temperature and returns 0 for v_turb.
 * alpha < 0 turns it off (explicit v_turb, the default). */
int rl_set_turbulence_alpha(rl_ctx *ctx, double alpha, double gamma, double mu);
/* Keplerian velocity about a star of mass mstar [g] (_calc_velocity
 * radlite.py:1634-1677, mstar from starinfo.inp): per cell v_r = v_theta
 * = 0, v_phi = sqrt(G mstar / r) [cm/s] with r the cell's radius, formed
 * on the device in place of rl_set_medium's velocity (trajectories are
 * rebuilt on the next trace).  Call after rl_set_medium. */
int rl_set_velocity_kepler(rl_ctx *ctx, double mstar);
/* Per cell [nr][nth]: gas temperature T_gas [K], dust temperature
 * T_dust [nspec][nr][nth] (NULL: T_gas for every species), log-abundance
 * (NULL: keep the abundance of rl_set_medium), turbulent velocity
 * v_turb [km/s]; molecular mass m_mol [m_p]; partition function Q = T/q
 * (q > 0), populations renormalised to sum 1 (q = 0), or the table of
 * rl_set_partition_table (q < 0).  Sets on the device: n_l = g_l
 * exp(-E_l/T)/Q (values below 1e-99 are 0), lw = sqrt(v_turb^2 +
 * 2kT/(m_mol m_p)) (or the alpha model above), abundance =
 * exp(log-abundance), j_d = sum_s rho_s kabs_s B(nu0, T_s), a_d = sum_s
 * rho_s (kabs_s + kscat_s), and aksmax = max lw.  k and m_p are
 * pyradlite's (astropy's CODATA 2018 values). */
int rl_set_physics(rl_ctx *ctx, const double *T_gas, const double *T_dust,
                   const double *log_abund, const double *v_turb_kms, double m_mol,
                   double q_over);
/* dL/d(T_gas, T_dust, log-abundance, v_turb) [nr][nth] (T_dust: [nspec]
 * [nr][nth], only when T_dust was given) from the cell gradients of the
 * last rl_trace_lines_adjoint call, at the fields of the last
 * rl_set_physics.  Any output may be NULL. */
int rl_physics_adjoint(rl_ctx *ctx, double *g_T_gas, double *g_T_dust,
                       double *g_log_abund, double *g_v_turb);
/* The device tables (populations [nr][nth][nlev], lw, abundance [nr][nth],
 * j_d, a_d [nr][nth][nlines]) to the host; any may be NULL. */
int rl_get_tables(rl_ctx *ctx, double *pops, double *lw, double *abund, double *jd,
                  double *ad);

/* ---- host pre-pass with the reference's arithmetic (R13) ---------- */
/* nu0 = 1.509160e26 * 1.986468498e-16 * (E_up - E_low)[cm^-1]. */
int rl_prepare_line_freqs(int nlines, const double *E_cm, const int *lev_up,
                          const int *lev_low, double *nu0);
/* grey dust j, alpha at each line centre from dust density/temperature
 * [nspec][nr][nth] and kappa_abs/kappa_scat [nspec][nfreq]. */
int rl_prepare_line_dust(int nr, int nth_half, int nspec, const double *rho_d,
                         const double *t_d, int nfreq, const double *cont_nu,
                         const double *kabs, const double *kscat, int nlines,
                         const double *nu0, double *jd, double *ad);
/* rl_prepare_line_dust with the dust scattering/PAH source scat
 * [nfreq][nr][nth] on the continuum grid (scatsource.dat plus
 * quantsource.dat, read_scatsource source.F:494-546; NULL: none), added
 * to j at each line centre as line.F:3593-3594 does. */
int rl_prepare_line_dust_scat(int nr, int nth_half, int nspec, const double *rho_d,
                              const double *t_d, int nfreq, const double *cont_nu,
                              const double *kabs, const double *kscat, const double *scat,
                              int nlines, const double *nu0, double *jd, double *ad);
/* starspectrum.inp F_nu at 1 pc -> I_nu at the stellar surface. */
int rl_prepare_star(int nfreq, const double *fnu_1pc, double rstar,
                    double *istar);
/* aksmax over cells (lw <= 0 falls back to fix_lw). */
double rl_prepare_aksmax(int ncell, const double *lw_kms, double fix_lw);
/* nfr = int(2 vmax/dv + 1), vmax := (nfr-1) dv / 2 (main.F:209-210). */
int rl_passband(double vmax, double dv, int *nfr, double *vmax_out);

#ifdef __cplusplus
}
#endif
#endif
