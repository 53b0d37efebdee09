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
constexpr int NF_RESET = 1;    // hole chord: ds := 0, init := 1 (telescope.F:4095-4104)
constexpr int NF_BLEND = 2;    // central-beam star blend here (telescope.F:4144-4193)
constexpr int NF_STAR = 4;     // I := I_star here: a rectangular-image ray (rbeam0 = 0)
                               // within R*, node and its predecessor in radial cell 1
                               // (telescope.F:4194-4208)

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

// hunt_bisect over the table xx(i) = tb[i], i = 1..n, from a guess g
// (0 <= g <= n) near the answer when use.  With P(j) = (x > xx(j)) ==
// ascnd, P(0) true and P(n+1) false, the bisection keeps P(jlo) and !P(jhi)
// and returns the jlo of the pair jhi = jlo + 1.  On a monotone table
// (ascending or descending) P is true up to one index and false after it,
// so that pair is unique: a candidate j with P(j) and !P(j+1) is the
// bisection's answer.  g-1, g, g+1 are tried, their loads at clamped
// indices so that they issue together without branches; else (or without
// use) the bisection runs.
__device__ inline int hunt_table(const double *__restrict__ tb, int n, double x, int g,
                                 bool use) {
    if (use) {
        const bool ascnd = tb[n] > tb[1];
        auto P = [&](int j) {
            const double v = tb[j < 1 ? 1 : (j > n ? n : j)];
            return j < 1 || (j <= n && (x > v) == ascnd);
        };
        const bool p0 = P(g - 1), p1 = P(g), p2 = P(g + 1), p3 = P(g + 2);
        if (p1 && !p2) return g;
        if (p0 && !p1) return g - 1;
        if (p2 && !p3) return g + 1;
    }
    return hunt_bisect([&](int i) { return tb[i]; }, n, x);
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
struct Carry { T phi0, srcl0, alpl0, S0; int init; T tau = T(0); };   // tau: char_tau

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
__device__ inline T qdr(T I, T js1, T alp1, T js2, T alp2, T ds, int &ovs, T &S0,
                        bool have_s0) {
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
    // js1/alp1 equals the previous element's js2/alp2 bit for bit when the
    // carry is valid, so it is reused rather than divided again
    T q2 = alp2 > T(0) ? ddiv<T>(js2, alp2) : T(0);
    T q1 = have_s0 ? S0 : (alp1 > T(0) ? ddiv<T>(js1, alp1) : T(0));
    T src1 = alp1 > T(0) ? q1 : (alp2 > T(0) ? q2 : T(0));
    T src2 = alp2 > T(0) ? q2 : (alp1 > T(0) ? q1 : T(0));
    S0 = q2;
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
    const bool have = c.init == 0;
    I = qdr<T>(I, src0, alp0, src1, alp1, ds, ovs, c.S0, have);
    c.tau = c.tau + T(0.5) * (alp0 + alp1) * ds;               // line.F:4598
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
    // the last (or only) element has one call site: a segment integrated
    // whole keeps a = p0, sa = 0 (ds - 0 = ds exactly) and its node widths
    Pt<T> a = p0;
    T sa = T(0), lwa = lw0, lwb = lw1;
    if (T(2.0) * crit * x > T(1.0)) {
        T s_c = ddiv<T>(ds * (L.uch - p0.u), p1.u - p0.u);
        T dls = ddiv<T>(ds, x);
        T sright = s_c + crit * dls, sleft = s_c - crit * dls;
        if (sright > T(0) && sleft < ds) {
            T lg_ds = ddiv<T>(sright - sleft, T(31 - 1.0));
            lwa = lwb = lw;
            // Only sub-points with 0 < s < ds are used; visiting the index
            // range that can contain them (with one index of margin on
            // each side) and keeping the exact test gives the same points.
            int ilo = 1, ihi = 31;
            if (lg_ds > T(0)) {
                T f0 = -sleft / lg_ds, f1 = (ds - sleft) / lg_ds;
                if (f0 > T(0) && f0 < T(40)) ilo = max(1, (int)f0);
                if (f1 > T(-10) && f1 < T(40)) ihi = min(31, (int)f1 + 2);
            }
            for (int i = ilo; i <= ihi; i++) {
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
        }
    }
    integ_elem<T>(L, I, ds - sa, a, p1, lwa, lwb, c, ovs);
}

// ---------------------------------------------------------------------
// Scaled-record formulation (the FP32 performance mode, and the
// arithmetic of the adjoint in both precisions).  Everything that does
// not depend on the channel is precomputed per node / per segment
// (FastRecT), and the quadrature weights of qdr_src_2 are evaluated
// without the cancellation in e1 = dtau - (1 - exp(-dtau)), which loses
// all single-precision digits for small dtau.
// ---------------------------------------------------------------------
template <typename T>
struct alignas(16 * sizeof(T) >= 64 ? 64 : 16) FastRecT {
    // node k: emissivities and opacities are stored divided by a per-node
    // scale sig (the dust opacity where there is dust), so that j/alpha is
    // formed from O(1) numbers even where j and alpha themselves are below
    // the single-precision range.  j = sig*(jdh + jch*phi), alpha =
    // sig*(adh + dch*phi) with phi the line profile.
    T ds;      // s_k - s_{k-1}
    T u;       // Omega.v/c
    T sig;     // scale (may underflow to 0 in float; never used as divisor)
    T jdh, adh;// dust j, alpha / sig
    T jch;     // 5.27296241956e-28 nu0 N_up A / sig         (line.F:4578)
    T dch;     // 5.27296241956e-28 nu0 (N_low B_lu - N_up B_ul) / sig
    T flag;    // NF_* bits
    // segment (k-1 -> k)
    T inv_aa;  // 1 / (3.33567e-6 |nu0| lwav)                 (line.F:2301)
    T norm;    // 0.56419583546 * inv_aa
    T dls;     // sub-grid line length ds/x, < 0 when no sub-grid
    T inv_du;  // 1 / (u_k - u_{k-1})
    T rel0, rel1;  // sig_{k-1}/sig_max, sig_k/sig_max (sub-grid weights)
    T sigm;    // sig_max = max(sig_{k-1}, sig_k)
    T pad;
};
using FastRec = FastRecT<float>;

template <typename T> struct FOps;
template <> struct FOps<float> {
    static __device__ inline float ex(float x) { return __expf(x); }
    // a/b through the hardware reciprocal (1 ulp) for well-scaled b
    static __device__ inline float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    // fused multiply-add: the single-precision path is not held to the
    // reference's rounding, so it contracts (the FP64 parity path does not)
    static __device__ inline float fma(float a, float b, float c) {
        return __builtin_fmaf(a, b, c);
    }
    // a source function j/alpha (b > 0) through the reciprocal; where b is
    // so small that 1/b would overflow or lose the denormal (line wings of
    // dust-free cells) both are first scaled by 2^64 (exact), without a
    // branch
    static __device__ inline float sdiv(float a, float b) {
        const float sc = b < 1.0e-30f ? 0x1p64f : 1.0f;
        return (a * sc) * __builtin_amdgcn_rcpf(b * sc);
    }
    static __device__ inline float fmn(float a, float b) { return fminf(a, b); }
    static __device__ inline float fmx(float a, float b) { return fmaxf(a, b); }
};
template <> struct FOps<double> {
    static __device__ inline double ex(double x) { return exp(x); }
    static __device__ inline double fdiv(double a, double b) { return a / b; }
    static __device__ inline double sdiv(double a, double b) { return a / b; }
    static __device__ inline double fma(double a, double b, double c) { return a * b + c; }
    static __device__ inline double fmn(double a, double b) { return fmin(a, b); }
    static __device__ inline double fmx(double a, double b) { return fmax(a, b); }
};

// qdr_src_2 weights: I_out = I*xp + a*S1 + b*S2.  ra = a/x (needed by the
// adjoint: db/dx = a/x, da/dx = xp - a/x).
template <typename T>
__device__ inline void qdr_weights_f(T x, T &a, T &b, T &xp, T &ra) {
    // single precision: the series also covers dtau <= 1e-6, where it and
    // the thin-element weights (a = b = dtau/2, xp = 1 - dtau) agree to
    // O(dtau^2), below float resolution; one branch fewer per element
    if (sizeof(T) == 4 || x > T(1.0e-6)) {
        if (x < T(0.5)) {
            // b = e1/x = sum_{n>=1} (-1)^{n+1} x^n/(n+1)!,  e0 = x (1 - b)
            T p;
            if (sizeof(T) == 4) {
                using O = FOps<T>;
                p = O::fma(x, T(-1) / 40320, T(1) / 5040);
                p = O::fma(x, -p, T(1) / 720);
                p = O::fma(x, -p, T(1) / 120);
                p = O::fma(x, -p, T(1) / 24);
                p = O::fma(x, -p, T(1) / 6);
                p = O::fma(x, -p, T(0.5));
            } else {
                p = T(0.5) - x * (T(1) / 6 - x * (T(1) / 24 - x * (T(1) / 120 -
                    x * (T(1) / 720 - x * (T(1) / 5040 - x * (T(1) / 40320))))));
            }
            if (sizeof(T) == 8)   // double: series continued to x^13/15!
                p = T(0.5) - x * (T(1) / 6 - x * (T(1) / 24 - x * (T(1) / 120 -
                    x * (T(1) / 720 - x * (T(1) / 5040 - x * (T(1) / 40320 -
                    x * (T(1) / 362880 - x * (T(1) / 3628800 - x * (T(1) / 39916800 -
                    x * (T(1) / 479001600 - x * (T(1) / 6227020800.0 -
                    x * (T(1) / 87178291200.0 - x * (T(1) / 1307674368000.0)))))))))))));
            b = x * p;
            T e0 = x * (T(1) - b);
            xp = T(1) - e0;
            a = e0 - b;
            ra = T(1) - b - p;
        } else {
            xp = FOps<T>::ex(-x);
            T e0 = T(1) - xp;
            b = FOps<T>::fdiv(x - e0, x);
            a = e0 - b;
            ra = FOps<T>::fdiv(a, x);
        }
    } else {
        a = T(0.5) * x;
        b = T(0.5) * x;
        xp = T(1) - x;
        ra = T(0.5);
    }
}

template <typename T> struct FastLineT { T nu0, dnu, uch; };
using FastLine = FastLineT<float>;

template <typename T> struct FPtT { T u, sig, jd, ad, jc, dc; };   // one point, scaled values
using FPt = FPtT<float>;

// What an element needs of its start point, carried from the element that
// ended there (the start point of an element is the end point of the
// previous one unless init is set): sa0 = sig*alpha/sig and sj0 = sig*j/sig
// there (the products the optical depth and the bracket add), ah0 =
// alpha/sig (its sign picks the source functions), S0 its source
// function, al0 its line alpha/sig (the maser test); u0 the velocity of
// the segment's start node (the sub-grid test); the optical depth along
// the ray, summed with a compensation term (Kahan).  Carrying these
// instead of the start point keeps the loop from copying six values per
// segment; the point is read again from its record where more of it is
// needed (init, sub-grid).
template <typename T> struct FastCarryT {
    T sa0 = T(0), sj0 = T(0), ah0 = T(0), S0 = T(0), al0 = T(0);
    int init = 1;
    T u0 = T(0), tau = T(0), tau_c = T(0);
};

// the carry of a start point p0 with line values jl0, al0 (/ sig)
template <typename T>
__device__ inline void carry_start(const FPtT<T> &p0, T jl0, T al0, FastCarryT<T> &c) {
    const T jh0 = p0.jd + jl0, ah0 = p0.ad + al0;
    c.sa0 = p0.sig * ah0;
    c.sj0 = p0.sig * jh0;
    c.ah0 = ah0;
    c.S0 = ah0 > T(0) ? FOps<T>::sdiv(jh0, ah0) : T(0);
    c.al0 = al0;
}

// integrate_element_linedust (line.F:4515-4624) on scaled values.  The
// start point is taken from the carry; start() supplies it (an FPtT) only
// when init is set.
template <typename T, typename Start>
__device__ inline void elem_core(const FastLineT<T> &L, T &I, T ds, const Start &start,
                                 const FPtT<T> &p1, T inv_aa, T norm, FastCarryT<T> &c,
                                 int &ovs) {
    using O = FOps<T>;
    if (c.init) {
        const FPtT<T> p0 = start();
        T z0 = O::fma(-L.nu0, p0.u, L.dnu) * inv_aa;
        T phi0 = norm * O::ex(-(z0 * z0));
        carry_start<T>(p0, p0.jc * phi0, p0.dc * phi0, c);
    }
    T z1 = O::fma(-L.nu0, p1.u, L.dnu) * inv_aa;
    T phi1 = norm * O::ex(-(z1 * z1));
    T jl1 = p1.jc * phi1, al1 = p1.dc * phi1;
    T jh1 = p1.jd + jl1;                          // j / sig
    T ah1 = p1.ad + al1;                          // alpha / sig
    const T ah0 = c.ah0;
    const T hds = T(0.5) * ds;
    T dtau = O::fma(p1.sig, ah1, c.sa0) * hds;
    T theomax = O::fma(p1.sig, jh1, c.sj0) * hds;
    {   // compensated sum: tau_c keeps the low-order part lost in tau
        const T y = dtau - c.tau_c, t = c.tau + y;
        c.tau_c = (t - c.tau) - y;
        c.tau = t;
    }
    T wa, wb, xp, ra;
    qdr_weights_f<T>(dtau, wa, wb, xp, ra);
    // source functions (transfer.F:1528-1541) from the scaled values
    T q1 = ah1 > T(0) ? FOps<T>::sdiv(jh1, ah1) : T(0);
    T q0 = c.S0;
    T s1 = ah0 > T(0) ? q0 : (ah1 > T(0) ? q1 : T(0));
    T s2 = ah1 > T(0) ? q1 : (ah0 > T(0) ? q0 : T(0));
    T q = dtau > T(1e-9f) ? O::fma(wa, s1, wb * s2) : theomax;
    ovs += q > T(2) * theomax;
    q = O::fmn(q, theomax);
    I = O::fma(I, xp, q);
    c.sa0 = p1.sig * ah1;
    c.sj0 = p1.sig * jh1;
    c.ah0 = ah1;
    c.S0 = q1;
    c.al0 = al1;
    c.init = 0;
}

template <typename T>
__device__ inline void elem_fast(const FastLineT<T> &L, T &I, T ds, const FPtT<T> &p0,
                                 const FPtT<T> &p1, T inv_aa, T norm, FastCarryT<T> &c,
                                 int &ovs) {
    elem_core<T>(L, I, ds, [&]() { return p0; }, p1, inv_aa, norm, c, ovs);
}

// Sub-grid points of segment k (clever_integrate_element_linedust,
// line.F:4706-4800): s_i = sleft + (i-1) lg for i in [ilo, ihi], used when
// 0 < s_i < ds.  Returns false when the segment is integrated whole.
template <typename T>
struct SubGrid { T sleft, lg; int ilo, ihi; };

template <typename T>
__device__ inline bool subgrid_of(const FastRecT<T> &b, T ds, T u0, T uch, SubGrid<T> &sg) {
    if (!(b.dls >= T(0))) return false;            // 6x > 1 (line.F:4715)
    T s_c = ds * (uch - u0) * b.inv_du;
    T sright = s_c + T(3) * b.dls, sleft = s_c - T(3) * b.dls;
    if (!(sright > T(0) && sleft < ds)) return false;
    T lg = (sright - sleft) * (T(1) / T(30));
    // only bounds of the index range (the exact 0 < s < ds test follows)
    T f0 = FOps<T>::fdiv(-sleft, lg), f1 = FOps<T>::fdiv(ds - sleft, lg);
    sg.sleft = sleft;
    sg.lg = lg;
    sg.ilo = f0 > T(0) ? max(1, (int)FOps<T>::fmn(f0, T(40))) : 1;
    sg.ihi = f1 < T(40) ? min(31, (int)FOps<T>::fmx(f1, T(-2)) + 2) : 31;
    return true;
}

// a sub-point at fraction e of the segment: linear interpolation of the
// unscaled quantities (line.F:4775-4790), done with relative weights
// (w receives the normalised weights n0, n1 and 1/W, for the adjoint)
template <typename T> struct SubW { T e, n0, n1, iw; };

template <typename T>
__device__ inline FPtT<T> subpoint_w(const FastRecT<T> &b, const FPtT<T> &p0,
                                     const FPtT<T> &p1, T e, SubW<T> &w) {
    T ep = T(1) - e;
    T w0 = ep * b.rel0, w1 = e * b.rel1, W = w0 + w1;
    T iw = FOps<T>::fdiv(T(1), W);
    w0 *= iw; w1 *= iw;
    w = SubW<T>{e, w0, w1, iw};
    return FPtT<T>{ep * p0.u + e * p1.u, b.sigm * W, w0 * p0.jd + w1 * p1.jd,
                   w0 * p0.ad + w1 * p1.ad, w0 * p0.jc + w1 * p1.jc, w0 * p0.dc + w1 * p1.dc};
}

template <typename T>
__device__ inline FPtT<T> subpoint(const FastRecT<T> &b, const FPtT<T> &p0, const FPtT<T> &p1,
                                   T e) {
    SubW<T> w;
    return subpoint_w<T>(b, p0, p1, e, w);
}

template <typename T>
__device__ inline FPtT<T> point_of(const FastRecT<T> &a) {
    return FPtT<T>{a.u, a.sig, a.jdh, a.adh, a.jch, a.dch};
}

// one segment (nodes k-1 -> k) of charintline on FastRecT data; prev is
// the record of node k-1, read only for a flagged or sub-gridded segment
// (the carry holds what a whole segment needs of its start)
template <typename T>
__device__ inline void segment_fast(const FastRecT<T> &b, const FastRecT<T> *prev,
                                    const FPtT<T> &p1, const FastLineT<T> &L, T &I, T istar,
                                    T starfract, FastCarryT<T> &c, int &ovs) {
    T ds = b.ds;
    const int fl = (int)b.flag;
    // flagged nodes are rare: the wave skips their handling unless one of
    // its lanes has one (a reset segment's ds is already 0 in the record)
    if (__builtin_amdgcn_ballot_w64(fl != 0)) {
        if (fl & NF_RESET) { ds = T(0); c.init = 1; }
        if (fl & NF_BLEND) { I = (T(1) - starfract) * I + starfract * istar; c.init = 1; }
        if (fl & NF_STAR) { I = istar; c.init = 1; }
    }
    // one call for the last (or only) element of every segment: a segment
    // integrated whole has sa = 0 and ds - 0 = ds exactly, and its start
    // point is node k-1; after a sub-element init is 0, so the start point
    // is not read
    SubGrid<T> sg;
    T sa = T(0);
    if (subgrid_of<T>(b, ds, c.u0, L.uch, sg)) {
        const FPtT<T> p0 = point_of<T>(*prev);
        FPtT<T> pp = p0;
        const T inv_ds = FOps<T>::fdiv(T(1), ds);
        for (int i = sg.ilo; i <= sg.ihi; i++) {
            T s = sg.sleft + (T)(i - 1) * sg.lg;
            if (s > T(0) && s < ds) {
                FPtT<T> q = subpoint<T>(b, p0, p1, s * inv_ds);
                elem_fast<T>(L, I, s - sa, pp, q, b.inv_aa, b.norm, c, ovs);
                pp = q;
                sa = s;
            }
        }
    }
    elem_core<T>(L, I, ds - sa, [&]() { return point_of<T>(*prev); }, p1, b.inv_aa, b.norm, c,
                 ovs);
    c.u0 = p1.u;
}

// charintline (telescope.F:3889-4312) on FastRecT data
template <typename T>
__device__ inline T trace_fast(const FastRecT<T> *__restrict__ rec, int n, const FastLineT<T> &L,
                               T I0, T istar, T starfract, int &ovs, int &maser,
                               T *tau = nullptr) {
    T I = I0;
    FastCarryT<T> c;
    c.u0 = rec[0].u;
    for (int k = 1; k < n; k++) {
        const FastRecT<T> b = rec[k];
        segment_fast<T>(b, rec + k - 1, point_of<T>(b), L, I, istar, starfract, c, ovs);
        if (b.sig * c.al0 * b.ds < T(-0.01f)) maser = 1;     // b.ds is 0 on a reset
    }
    if (tau) *tau = c.tau;
    return I;
}

// Node record produced per (item, node): the channel-independent values
// of get_line_dust_values (line.F:3965-4217) plus the segment length.
template <typename T>
struct alignas(sizeof(T) * 8) NodeRec {
    T ds, u, jd, ad, lw, nu, nl, flag;
};

}  // namespace rl
