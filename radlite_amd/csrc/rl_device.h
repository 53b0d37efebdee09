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
struct Carry { T phi0, srcl0, alpl0, S0; int init; };

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
            integ_elem<T>(L, I, ds - sa, a, p1, lw, lw, c, ovs);
            return;
        }
    }
    integ_elem<T>(L, I, ds, p0, p1, lw0, lw1, c, ovs);
}

// ---------------------------------------------------------------------
// FP32 performance formulation.  Everything that does not depend on the
// channel is precomputed per node / per segment (FastRec), and the
// quadrature weights of qdr_src_2 are evaluated without the cancellation
// in e1 = dtau - (1 - exp(-dtau)), which loses all single-precision
// digits for small dtau.
// ---------------------------------------------------------------------
struct alignas(16) FastRec {
    // node k: emissivities and opacities are stored divided by a per-node
    // scale sig (the dust opacity where there is dust), so that j/alpha is
    // formed from O(1) numbers even where j and alpha themselves are below
    // the single-precision range.  j = sig*(jdh + jch*phi), alpha =
    // sig*(adh + dch*phi) with phi the line profile.
    float ds;      // s_k - s_{k-1}
    float u;       // Omega.v/c
    float sig;     // scale (may underflow to 0 in float; never used as divisor)
    float jdh, adh;// dust j, alpha / sig
    float jch;     // 5.27296241956e-28 nu0 N_up A / sig         (line.F:4578)
    float dch;     // 5.27296241956e-28 nu0 (N_low B_lu - N_up B_ul) / sig
    float flag;    // NF_* bits
    // segment (k-1 -> k)
    float inv_aa;  // 1 / (3.33567e-6 |nu0| lwav)                 (line.F:2301)
    float norm;    // 0.56419583546 * inv_aa
    float dls;     // sub-grid line length ds/x, < 0 when no sub-grid
    float inv_du;  // 1 / (u_k - u_{k-1})
    float rel0, rel1;  // sig_{k-1}/sig_max, sig_k/sig_max (sub-grid weights)
    float sigm;    // sig_max = max(sig_{k-1}, sig_k)
    float pad;
};

// qdr_src_2 weights: I_out = I*xp + a*S1 + b*S2
__device__ inline void qdr_weights_f(float x, float &a, float &b, float &xp) {
    if (x > 1.0e-6f) {
        if (x < 0.5f) {
            // b = e1/x = sum_{n>=1} (-1)^{n+1} x^n/(n+1)!,  e0 = x (1 - b)
            b = x * (0.5f - x * (1.0f / 6 - x * (1.0f / 24 - x * (1.0f / 120 -
                x * (1.0f / 720 - x * (1.0f / 5040 - x * (1.0f / 40320)))))));
            float e0 = x * (1.0f - b);
            xp = 1.0f - e0;
            a = e0 - b;
        } else {
            xp = __expf(-x);
            float e0 = 1.0f - xp;
            b = __fdividef(x - e0, x);
            a = e0 - b;
        }
    } else {
        a = 0.5f * x;
        b = 0.5f * x;
        xp = 1.0f - x;
    }
}

struct FastLine { float nu0, dnu, uch; };

struct FPt { float u, sig, jd, ad, jc, dc; };     // one point, scaled values

struct FastCarry { float jl0, al0, S0; int init; };  // line j, alpha / sig at the start

// integrate_element_linedust (line.F:4515-4624) on scaled values
__device__ inline void elem_fast(const FastLine &L, float &I, float ds, const FPt &p0,
                                 const FPt &p1, float inv_aa, float norm, FastCarry &c,
                                 int &ovs) {
    if (c.init) {
        float z0 = (L.dnu - L.nu0 * p0.u) * inv_aa;
        float phi0 = norm * __expf(-(z0 * z0));
        c.jl0 = p0.jc * phi0;
        c.al0 = p0.dc * phi0;
    }
    float z1 = (L.dnu - L.nu0 * p1.u) * inv_aa;
    float phi1 = norm * __expf(-(z1 * z1));
    float jl1 = p1.jc * phi1, al1 = p1.dc * phi1;
    float jh0 = p0.jd + c.jl0, jh1 = p1.jd + jl1;      // j / sig
    float ah0 = p0.ad + c.al0, ah1 = p1.ad + al1;      // alpha / sig
    float a0 = p0.sig * ah0, a1 = p1.sig * ah1;
    float dtau = 0.5f * (a0 + a1) * ds;
    float theomax = 0.5f * (p0.sig * jh0 + p1.sig * jh1) * ds;
    float wa, wb, xp;
    qdr_weights_f(dtau, wa, wb, xp);
    // source functions (transfer.F:1528-1541) from the scaled values; the
    // start point equals the previous end point unless init is set
    float q1 = ah1 > 0.f ? jh1 / ah1 : 0.f;
    float q0 = c.init ? (ah0 > 0.f ? jh0 / ah0 : 0.f) : c.S0;
    float s1 = ah0 > 0.f ? q0 : (ah1 > 0.f ? q1 : 0.f);
    float s2 = ah1 > 0.f ? q1 : (ah0 > 0.f ? q0 : 0.f);
    float q = dtau > 1e-9f ? wa * s1 + wb * s2 : theomax;
    ovs += q > 2.f * theomax;
    q = fminf(q, theomax);
    I = I * xp + q;
    c.jl0 = jl1;
    c.al0 = al1;
    c.S0 = q1;
    c.init = 0;
}

// charintline (telescope.F:3889-4312) on FastRec data
__device__ inline float trace_fast(const FastRec *__restrict__ rec, int n, const FastLine &L,
                                   float I0, float istar, float starfract, int &ovs,
                                   int &maser) {
    float I = I0;
    FastCarry c{0.f, 0.f, 0.f, 1};
    const FastRec a = rec[0];
    FPt p0{a.u, a.sig, a.jdh, a.adh, a.jch, a.dch};
    for (int k = 1; k < n; k++) {
        const FastRec b = rec[k];
        const FPt p1{b.u, b.sig, b.jdh, b.adh, b.jch, b.dch};
        float ds = b.ds;
        const int fl = (int)b.flag;
        if (fl & NF_RESET) { ds = 0.f; c.init = 1; }
        if (fl & NF_BLEND) { I = (1.f - starfract) * I + starfract * istar; c.init = 1; }
        bool done = false;
        if (b.dls >= 0.f) {                  // 6x > 1 (line.F:4715)
            float s_c = ds * (L.uch - p0.u) * b.inv_du;
            float sright = s_c + 3.f * b.dls, sleft = s_c - 3.f * b.dls;
            if (sright > 0.f && sleft < ds) {
                float lg = (sright - sleft) * (1.f / 30.f);
                float f0 = -sleft / lg, f1 = (ds - sleft) / lg;
                int ilo = f0 > 0.f ? max(1, (int)fminf(f0, 40.f)) : 1;
                int ihi = f1 < 40.f ? min(31, (int)fmaxf(f1, -2.f) + 2) : 31;
                FPt pp = p0;
                float sa = 0.f;
                const float inv_ds = 1.f / ds;
                for (int i = ilo; i <= ihi; i++) {
                    float s = sleft + (float)(i - 1) * lg;
                    if (s > 0.f && s < ds) {
                        // linear interpolation of the unscaled quantities
                        // (line.F:4775-4790), done with relative weights
                        float e = s * inv_ds, ep = 1.f - e;
                        float w0 = ep * b.rel0, w1 = e * b.rel1, W = w0 + w1;
                        float iw = 1.f / W;
                        w0 *= iw; w1 *= iw;
                        FPt q{ep * p0.u + e * p1.u, b.sigm * W, w0 * p0.jd + w1 * p1.jd,
                              w0 * p0.ad + w1 * p1.ad, w0 * p0.jc + w1 * p1.jc,
                              w0 * p0.dc + w1 * p1.dc};
                        elem_fast(L, I, s - sa, pp, q, b.inv_aa, b.norm, c, ovs);
                        pp = q;
                        sa = s;
                    }
                }
                elem_fast(L, I, ds - sa, pp, p1, b.inv_aa, b.norm, c, ovs);
                done = true;
            }
        }
        if (!done) elem_fast(L, I, ds, p0, p1, b.inv_aa, b.norm, c, ovs);
        if (b.sig * c.al0 * ds < -0.01f) maser = 1;
        p0 = p1;
    }
    return I;
}

// Node record produced per (item, node): the channel-independent values
// of get_line_dust_values (line.F:3965-4217) plus the segment length.
template <typename T>
struct alignas(sizeof(T) * 8) NodeRec {
    T ds, u, jd, ad, lw, nu, nl, flag;
};

}  // namespace rl
