// The per-cell pre-pass on the device, and its adjoint.
//
// pyradlite forms the inputs of a RADLite run from the physical model
// (radlite.py:1024-1127 _prep_mol_forcore, 1592-1631 _calc_turbulence,
// 1481-1531 _calc_abundance) and RADLite adds the dust at each line centre
// (line_dust_compute_src_alp, line.F:3560-3607; bplanck, setup.F:937-952).
// Here those tables are computed where the trace reads them, from per-cell
// temperature, log-abundance and turbulent velocity, so that a fit loop
// moves three numbers per cell across PCIe instead of the level
// populations and the per-line dust tables; the adjoint maps the trace's
// cell gradients (left on the device by the trace adjoint) back to those
// three fields.
//
//   populations  n_l = g_l exp(-E_l/T) / (T/q)            (q > 0)
//                n_l = g_l exp(-E_l/T) / sum_m g_m exp(-E_m/T)   (q = 0)
//                n_l = g_l exp(-E_l/T) / Q(T), Q linear in a tabulated
//                partition sum and extrapolated past its ends (q < 0:
//                scipy's interp1d(kind="linear", fill_value="extrapolate")
//                as radlite.py:1150-1153 applies it to ParSum.dat)
//                values below 1e-99 are 0 (radlite.py:1127)
//   width        lw = sqrt(v_t^2 + 2 k T / (m_mol m_p)) [km/s], or the
//                alpha model of _calc_turbulence (radlite.py:1613-1622):
//                lw = sqrt((alpha c_s)^2 + 2 k T_d / (m_mol m_p)), c_s^2 =
//                gamma k T_d / (mu m_p), T_d the first dust temperature
//   velocity     Keplerian (radlite.py:1653-1664): v = (0, 0, sqrt(G M / r))
//                with r the cell's radius
//   abundance    exp(log_abund)
//   dust         j_d = sum_s rho_s kabs_s(nu0) B(nu0, T_s) (+ the scattering
//                source at nu0 when one is set) ; a_d = sum_s rho_s
//                (kabs_s + kscat_s), with B = 1.47455e-47 nu^3 /
//                (exp(4.7989e-11 nu/T) - 1) + 1e-290 (0 at T = 0)
//
// Tables are [cell][level] and [cell][line], cells ir-major with theta
// fastest, as rl_set_levels / rl_set_dust_line take them.
#include "rl_engine.h"

#include <hip/hip_runtime.h>

namespace rl {

constexpr double PRE_BA = 1.47455e-47, PRE_BX = 4.7989e-11;

__device__ inline double wave_sum_d(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The partition sum at t from a table (x ascending, n >= 2), as scipy's
// interp1d(kind="linear", bounds_error=False, fill_value="extrapolate")
// evaluates it: the interval is searchsorted(x, t) (the first x >= t)
// clipped to [1, n-1], and Q = slope (t - x_lo) + y_lo with slope = (y_hi -
// y_lo) / (x_hi - x_lo); dQ is that slope.
__device__ inline double q_table(double t, int n, const double *__restrict__ x,
                                 const double *__restrict__ y, double &dQ) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (x[m] < t) lo = m + 1; else hi = m;
    }
    const int i = min(max(lo, 1), n - 1);
    dQ = (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    return dQ * (t - x[i - 1]) + y[i - 1];
}

// one wave per cell: populations over the levels (lanes), width, abundance
__global__ void __launch_bounds__(64)
k_pre_cells(int ncell, int nlev, const double *__restrict__ T, const double *__restrict__ la,
            const double *__restrict__ vt, const double *__restrict__ Tw,
            const double *__restrict__ EK, const double *__restrict__ g, double q, int nq,
            const double *__restrict__ qx, const double *__restrict__ qy, double cw,
            double *__restrict__ pops, double *__restrict__ lw, double *__restrict__ abund) {
    const int c = blockIdx.x;
    if (c >= ncell) return;
    const int lane = threadIdx.x;
    const double t = T[c];
    double *p = pops + (size_t)c * nlev;
    if (q > 0.0 || q < 0.0) {
        double dQ;
        const double Q = q > 0.0 ? t / q : q_table(t, nq, qx, qy, dQ);
        for (int l = lane; l < nlev; l += 64) {
            const double n = g[l] * exp(-EK[l] / t) / Q;
            p[l] = n < 1e-99 ? 0.0 : n;
        }
    } else {
        double s = 0.0;
        for (int l = lane; l < nlev; l += 64) s += g[l] * exp(-EK[l] / t);
        s = wave_sum_d(s);
        for (int l = lane; l < nlev; l += 64) {
            const double n = g[l] * exp(-EK[l] / t) / s;
            p[l] = n < 1e-99 ? 0.0 : n;
        }
    }
    if (lane == 0) {
        // explicit turbulence: v_t^2 + cw T; the alpha model: (cw + ca) T_d
        // with cw carrying both coefficients and no v_t
        lw[c] = vt ? sqrt(vt[c] * vt[c] + cw * t) : sqrt(cw * Tw[c]);
        if (la) abund[c] = exp(la[c]);
    }
}

// Keplerian velocity about the star (radlite.py:1653-1664): per cell (ir
// major, theta fastest) (v_r, v_theta, v_phi) = (0, 0, sqrt(G M / r_ir))
__global__ void k_pre_kepler(int ncell, int nh, const double *__restrict__ r, double gm,
                             double *__restrict__ vel) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const double v = sqrt(gm / r[c / nh]);
    vel[3 * (size_t)c + 0] = 0.0;
    vel[3 * (size_t)c + 1] = 0.0;
    vel[3 * (size_t)c + 2] = v;
}

__device__ inline double bplanck_d(double temp, double nu) {
    if (temp == 0.0) return 0.0;
    return PRE_BA * nu * nu * nu / (exp(PRE_BX * nu / temp) - 1.0) + 1.0e-290;
}

// dB/dT of the above (0 at T = 0)
__device__ inline double dbplanck_dt(double temp, double nu) {
    if (temp == 0.0) return 0.0;
    const double x = PRE_BX * nu / temp, e = exp(x), d = e - 1.0;
    return PRE_BA * nu * nu * nu * e / (d * d) * x / temp;
}

// one thread per (cell, line), line fastest: the dust at the line centre,
// summed over species in the reference's order (absorption with the
// emission, then scattering)
__global__ void k_pre_dust(long long n, int nl, int nspec, long long ncell,
                           const double *__restrict__ nu0, const double *__restrict__ rho_d,
                           const double *__restrict__ Td, const double *__restrict__ kabs,
                           const double *__restrict__ kscat, const double *__restrict__ scat,
                           const int *__restrict__ scat_lo, const double *__restrict__ scat_w,
                           int scat_nf, double *__restrict__ jd, double *__restrict__ ad) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long c = i / nl;
    const int l = (int)(i - c * nl);
    const double freq = nu0[l];
    double src = 0.0, alp = 0.0;
    for (int s = 0; s < nspec; s++) {
        const double kw = kabs[(size_t)s * nl + l], rd = rho_d[s * ncell + c];
        src = src + rd * kw * bplanck_d(Td[s * ncell + c], freq);
        alp = alp + rd * kw;
    }
    // scattering/PAH source at the line centre (line.F:3593-3594)
    if (scat) {
        const int lo = scat_lo[l];
        if (lo > 0 && lo < scat_nf) {
            const double w = scat_w[l];
            src = src + w * scat[(size_t)lo * ncell + c] + (1.0 - w) * scat[(size_t)(lo - 1) * ncell + c];
        }
    }
    for (int s = 0; s < nspec; s++) alp = alp + rho_d[s * ncell + c] * kscat[(size_t)s * nl + l];
    jd[i] = src;
    ad[i] = alp;
}

// block maxima of x (for aksmax = max lw)
__global__ void k_pre_max(int n, const double *__restrict__ x, double *__restrict__ part) {
    __shared__ double m[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    m[threadIdx.x] = i < n ? x[i] : -1.0e300;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = m[0];
}

// adjoint, one wave per cell: dL/dT from the populations, the width and
// (when the dust follows the gas temperature) the dust emission; dL/dv_t;
// dL/dlog_abund
__global__ void __launch_bounds__(64)
k_pre_cells_adj(int ncell, int nlev, int nl, int nspec, int dust_from_gas,
                const double *__restrict__ T, const double *__restrict__ vt,
                const double *__restrict__ Tw, const double *__restrict__ EK,
                const double *__restrict__ g, double q, int nq, const double *__restrict__ qx,
                const double *__restrict__ qy, double cw, const double *__restrict__ pops,
                const double *__restrict__ lw,
                const double *__restrict__ abund, const double *__restrict__ g_pops,
                const double *__restrict__ g_lw, const double *__restrict__ g_abund,
                const double *__restrict__ nu0, const double *__restrict__ rho_d,
                const double *__restrict__ Td, const double *__restrict__ kabs,
                const double *__restrict__ g_jd, double *__restrict__ gT,
                double *__restrict__ gTd, double *__restrict__ gvt, double *__restrict__ gla) {
    const int c = blockIdx.x;
    if (c >= ncell) return;
    const int lane = threadIdx.x;
    const double t = T[c];
    const double *p = pops + (size_t)c * nlev, *gp = g_pops + (size_t)c * nlev;
    double acc = 0.0;
    if (q > 0.0) {
        // n = g exp(-E/T) q / T: dn/dT = n (E/T^2 - 1/T); 0 where clamped to 0
        for (int l = lane; l < nlev; l += 64)
            acc += gp[l] * p[l] * (EK[l] / (t * t) - 1.0 / t);
    } else if (q < 0.0) {
        // n = g exp(-E/T) / Q(T): dn/dT = n (E/T^2 - Q'/Q), Q' the slope of
        // the table interval that forms Q
        double dQ;
        const double Q = q_table(t, nq, qx, qy, dQ);
        const double r = dQ / Q;
        for (int l = lane; l < nlev; l += 64)
            acc += gp[l] * p[l] * (EK[l] / (t * t) - r);
    } else {
        // n = w / sum w: dn/dT = n (E - <E>) / T^2
        double e = 0.0;
        for (int l = lane; l < nlev; l += 64) e += p[l] * EK[l];
        e = wave_sum_d(e);
        for (int l = lane; l < nlev; l += 64) acc += gp[l] * p[l] * (EK[l] - e) / (t * t);
    }
    // dust emission through the dust temperature(s)
    double accd = 0.0;
    const double *gj = g_jd + (size_t)c * nl;
    for (int s = 0; s < nspec; s++) {
        const double rd = rho_d[(size_t)s * ncell + c], td = Td[(size_t)s * ncell + c];
        double a = 0.0;
        for (int l = lane; l < nl; l += 64)
            a += gj[l] * rd * kabs[(size_t)s * nl + l] * dbplanck_dt(td, nu0[l]);
        a = wave_sum_d(a);
        if (dust_from_gas) accd += a;
        else if (lane == 0) gTd[(size_t)s * ncell + c] = a;
    }
    acc = wave_sum_d(acc);
    if (lane == 0) {
        const double w = lw[c];
        const double glw_t = g_lw[c] * (0.5 * cw / w);   // d lw / d (its temperature)
        if (vt) {
            gT[c] = acc + glw_t + accd;
            gvt[c] = g_lw[c] * (vt[c] / w);
        } else {
            // the alpha model's width follows the first dust temperature
            // (the gas temperature where the dust follows the gas)
            gT[c] = acc + accd + (dust_from_gas ? glw_t : 0.0);
            if (!dust_from_gas) gTd[c] += glw_t;
            gvt[c] = 0.0;
        }
        gla[c] = g_abund[c] * abund[c];
    }
}

void launch_prepass(hipStream_t st, const PrepassLaunch &a) {
    hipLaunchKernelGGL(k_pre_cells, dim3(a.ncell), dim3(64), 0, st, a.ncell, a.nlev, a.T, a.la,
                       a.vt, a.Tw, a.EK, a.g, a.q, a.nq, a.qx, a.qy, a.cw, a.pops, a.lw,
                       a.abund);
    const long long n = (long long)a.ncell * a.nlines;
    hipLaunchKernelGGL(k_pre_dust, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                       a.nlines, a.nspec, (long long)a.ncell, a.nu0, a.rho_d, a.Td, a.kabs,
                       a.kscat, a.scat, a.scat_lo, a.scat_w, a.scat_nf, a.jd, a.ad);
    hipLaunchKernelGGL(k_pre_max, dim3((a.ncell + 255) / 256), dim3(256), 0, st, a.ncell, a.lw,
                       a.part);
}

void launch_prepass_adjoint(hipStream_t st, const PrepassLaunch &a) {
    hipLaunchKernelGGL(k_pre_cells_adj, dim3(a.ncell), dim3(64), 0, st, a.ncell, a.nlev,
                       a.nlines, a.nspec, a.dust_from_gas, a.T, a.vt, a.Tw, a.EK, a.g, a.q,
                       a.nq, a.qx, a.qy, a.cw, a.pops, a.lw, a.abund, a.g_pops, a.g_lw, a.g_abund, a.nu0, a.rho_d, a.Td,
                       a.kabs, a.g_jd, a.gT, a.gTd, a.gvt, a.gla);
}

void launch_kepler(hipStream_t st, int ncell, int nh, const double *r, double gm, double *vel) {
    hipLaunchKernelGGL(k_pre_kepler, dim3((ncell + 255) / 256), dim3(256), 0, st, ncell, nh, r,
                       gm, vel);
}

}  // namespace rl
