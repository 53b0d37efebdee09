// Device-side building blocks of the raytracer (gfx950, wave64).
//
// Everything here is a restatement of the reference's per-segment
// physics in the reference's own evaluation order; the FP64
// instantiation is the parity mode, the FP32 instantiation the
// performance mode.  The translation unit is compiled with
// -ffp-contract=off so that FP64 arithmetic rounds exactly as the
// reference's (x86-64, no FMA) evaluation does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rl {

constexpr double TELESC_EPS = 1.0e-10;              // configure.h:46
constexpr double PICONST = 3.1415926535897932385;   // main.h:452
constexpr int RAYEXPT = 100;                        // common_telescope.h:4
constexpr int RAYADPT = 4;
constexpr int RAYRNPT = 4;
constexpr int EXMAX = 40;     // extra points per ray: at most 1+16+1+16
constexpr double LEVTHRES = 1.0e-3;                 // configure.h:52

// Error codes are the reference's stop codes.
enum : int { E_OK = 0, E_INPUT = 13, E_STAR = 124, E_DSNEG = 749,
             E_HOLE = 7454, E_DT = 6023, E_DR = 6024, E_HUNT = 192,
             E_EXEQ = 987, E_EXORD = 988, E_EXOVF = 83459, E_CAP = 155 };

__device__ inline void set_err(int *err, int code) {
    atomicCAS(err, 0, code);
}

// Numerical Recipes hunt as a pure bisection (nrecip.F:157-204 with an
// out-of-range guess): jlo with xx[jlo] < x <= xx[jlo+1], xx 1-based.
template <typename F>
__device__ inline int hunt_bisect(F xx, int n, double x) {
    bool ascnd = xx(n) > xx(1);
    int jlo = 0, jhi = n + 1;
    while (jhi - jlo != 1) {
        int jm = (jhi + jlo) >> 1;
        if ((x > xx(jm)) == ascnd) jlo = jm; else jhi = jm;
    }
    return jlo;
}

// make_index theta re-index (interpol.F:87-100)
__device__ inline int ridx_it(int it, int ny) {
    int v = it;
    if (v < 1) v = 1 - v;
    if (v > ny) v = 2 * ny + 1 - v;
    if (v > ny / 2) v = ny + 1 - v;
    return v;
}

// ---------------------------------------------------------------------
// Per-segment physics, templated on the arithmetic type.
// ---------------------------------------------------------------------
template <typename T> __device__ inline T dexp(T x);
template <> __device__ inline double dexp<double>(double x) { return exp(x); }
template <> __device__ inline float dexp<float>(float x) { return __expf(x); }
template <typename T> __device__ inline T ddiv(T a, T b);
template <> __device__ inline double ddiv<double>(double a, double b) { return a / b; }
template <> __device__ inline float ddiv<float>(float a, float b) { return __fdividef(a, b); }

template <typename T>
struct LineK {            // per (line, channel) constants
    T nu0;                // linefreq (signed) -- line.F:4554-4588
    T c_aa;               // 3.33567e-6*|nu0| (voigt_profile aa = c_aa*lw)
    T hknu;               // 5.27296241956e-28*nu0
    T A, Bud, Bdu;
    T dnu;                // line_dnu of this channel
    T uch;                // line_dnu/line_nu0 (telescope.F:392, line.F:4733)
};

template <typename T>
struct Carry { T phi0, srcl0, alpl0; int init; };

template <typename T>
struct Pt { T srcd, alpd, u, nu, nl; };     // one integration point

// voigt_profile, line.F:2280-2314 (pure Gaussian)
template <typename T>
__device__ inline T voigt(const LineK<T> &L, T aks, T dnu) {
    T aa = L.c_aa * aks;
    T norm = ddiv<T>(T(0.56419583546), aa);
    T u = ddiv<T>(dnu, aa);
    return norm * dexp<T>(-(u * u));
}

// qdr_src_2, transfer.F:1498-1571 -- I <- I*xp + min(Q, theomax)
template <typename T>
__device__ inline T qdr(T I, T js1, T alp1, T js2, T alp2, T ds, int &ovs) {
    T dtau1 = T(0.5) * (alp1 + alp2) * ds;
    T theomax = T(0.5) * (js1 + js2) * ds;
    T a, b, xp;
    if (dtau1 > T(1.0e-6)) {
        xp = dexp<T>(-dtau1);
        T e0 = T(1) - xp;
        T e1 = dtau1 - e0;
        b = ddiv<T>(e1, dtau1);
        a = e0 - b;
    } else {
        a = T(0.5) * dtau1;
        b = T(0.5) * dtau1;
        xp = T(1) - dtau1;
    }
    T src1, src2;
    if (alp1 > T(0)) src1 = ddiv<T>(js1, alp1);
    else if (alp2 > T(0)) src1 = ddiv<T>(js2, alp2);
    else src1 = T(0);
    if (alp2 > T(0)) src2 = ddiv<T>(js2, alp2);
    else if (alp1 > T(0)) src2 = ddiv<T>(js1, alp1);
    else src2 = T(0);
    T q;
    if (dtau1 > T((double)1e-9f)) q = a * src1 + b * src2;   // 1e-9 is REAL*4
    else q = T(0.5) * (js1 + js2) * ds;
    if (q > T(2) * theomax) ovs++;
    q = q < theomax ? q : theomax;
    return I * xp + q;
}

// integrate_element_linedust, line.F:4515-4624
template <typename T>
__device__ inline void integ_elem(const LineK<T> &L, T &I, T ds,
                                  const Pt<T> &p0, const Pt<T> &p1,
                                  T lw0, T lw1, Carry<T> &c, int &ovs) {
    T lwav = T(0.5) * (lw0 + lw1);
    T dnu1 = L.dnu - L.nu0 * p1.u;
    if (c.init == 1) {
        T dnu0 = L.dnu - L.nu0 * p0.u;
        c.phi0 = voigt<T>(L, lwav, dnu0);
        c.srcl0 = L.hknu * p0.nu * L.A * c.phi0;
        c.alpl0 = L.hknu * c.phi0 * (p0.nl * L.Bdu - p0.nu * L.Bud);
    }
    T phi1 = voigt<T>(L, lwav, dnu1);
    T srcl1 = L.hknu * p1.nu * L.A * phi1;
    T alpl1 = L.hknu * phi1 * (p1.nl * L.Bdu - p1.nu * L.Bud);
    T src0 = p0.srcd + c.srcl0, src1 = p1.srcd + srcl1;
    T alp0 = p0.alpd + c.alpl0, alp1 = p1.alpd + alpl1;
    I = qdr<T>(I, src0, alp0, src1, alp1, ds, ovs);
    c.phi0 = phi1;
    c.srcl0 = srcl1;
    c.alpl0 = alpl1;
    c.init = 0;
}

// clever_integrate_element_linedust, line.F:4636-4848 (SUBGRID on):
// sub-points are generated on the fly instead of stored.
template <typename T>
__device__ inline void clever_elem(const LineK<T> &L, T &I, T ds,
                                   const Pt<T> &p0, const Pt<T> &p1,
                                   T lw0, T lw1, Carry<T> &c, int &ovs) {
    const T crit = T(3.0);
    T lw = T(0.5) * (lw0 + lw1);
    T x = fabs(ddiv<T>(p1.u - p0.u, ddiv<T>(lw, T(2.99792458e5))));
    if (T(2.0) * crit * x > T(1.0)) {
        T s_c = ddiv<T>(ds * (L.uch - p0.u), p1.u - p0.u);
        T dls = ddiv<T>(ds, x);
        T sright = s_c + crit * dls, sleft = s_c - crit * dls;
        if (sright > T(0) && sleft < ds) {
            T lg_ds = ddiv<T>(sright - sleft, T(31 - 1.0));
            Pt<T> a = p0;
            T sa = T(0);
            for (int i = 1; i <= 31; i++) {
                T s = sleft + T(i - 1) * lg_ds;
                if (s > T(0) && s < ds) {
                    T e = ddiv<T>(s, ds), ep = T(1) - e;
                    Pt<T> b;
                    b.nu = ep * p0.nu + e * p1.nu;
                    b.nl = ep * p0.nl + e * p1.nl;
                    b.u = ep * p0.u + e * p1.u;
                    b.srcd = ep * p0.srcd + e * p1.srcd;
                    b.alpd = ep * p0.alpd + e * p1.alpd;
                    integ_elem<T>(L, I, s - sa, a, b, lw, lw, c, ovs);
                    a = b;
                    sa = s;
                }
            }
            integ_elem<T>(L, I, ds - sa, a, p1, lw, lw, c, ovs);
            return;
        }
    }
    integ_elem<T>(L, I, ds, p0, p1, lw0, lw1, c, ovs);
}

// Node record produced per (item, node): the channel-independent values
// of get_line_dust_values (line.F:3965-4217) plus the segment length.
template <typename T>
struct alignas(sizeof(T) * 8) NodeRec {
    T ds, u, jd, ad, lw, nu, nl, flag;
};
constexpr int NF_RESET = 1;    // hole chord: ds := 0, init := 1 (telescope.F:4095-4104)
constexpr int NF_BLEND = 2;    // central-beam star blend here (telescope.F:4144-4193)

}  // namespace rl
