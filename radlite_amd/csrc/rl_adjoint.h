// Reverse-mode (adjoint) formal solution on scaled records.
//
// The forward model per (ray, channel) is the recurrence of charintline
// (telescope.F:3889-4312): I_out = I_in * xp + min(Q, theomax) per
// element (qdr_src_2, transfer.F:1498-1571), elements being whole
// segments or the sub-grid steps of clever_integrate_element_linedust
// (line.F:4636-4848).  Its adjoint runs the segments backwards with
// lambda = dL/dI, taking I at each segment start from checkpoints kept
// by the forward pass (every segment, or every K-th with the ones between
// recomputed), and I at each sub-step start by recomputing the segment.  Everything that depends on the
// differentiated cell quantities is differentiated: the line profile
// through the width and the projected velocity of the points, the sub-grid
// point positions through the width, the source functions, the
// quadrature weights and the theomax bracket.
//
// Gradients are taken with respect to the scaled values of a point
// (j = sig*(jd + jc*phi), alpha = sig*(ad + dc*phi)); sig is a fixed
// per-node constant, so dL/d(true value) = dL/d(scaled value) / sig.
#pragma once
#include "rl_device.h"

namespace rl {

constexpr int ADJ_K = 16;        // largest checkpoint interval in segments

template <typename T>
struct PGrad { T u, sig, jd, ad, jc, dc; };

template <typename T>
__device__ inline void pg_zero(PGrad<T> &g) {
    g.u = g.sig = g.jd = g.ad = g.jc = g.dc = T(0);
}

// reverse of elem_fast.  Inputs: the element as the forward saw it (I_in,
// ds, end points, the end point's profile argument z1 and value phi1,
// profile constants, init and the carried line values jl0c/al0c used when
// init == 0), lam = dL/dI_out, and gjl1/gal1 = dL/d(line j, alpha at the
// end point) arriving from the next element, which carried them.  The
// forward quantities are recomputed with elem_fast's operations, so they
// are the forward's values.  Accumulates into g0/g1 (the two points),
// g_ds and g_lnaa (= dL/d ln inv_aa); returns dL/dI_in and, when
// init == 0, the gradient on the carried start values in gjl0/gal0.
template <typename T>
__device__ inline T elem_rev(const FastLineT<T> &L, T I_in, T lam, T ds, const FPtT<T> &p0,
                             const FPtT<T> &p1, T z1, T phi1, T inv_aa, T norm, int init,
                             T jl0c, T al0c, T gjl1, T gal1, PGrad<T> &g0, PGrad<T> &g1,
                             T &g_ds, T &g_lnaa, T &gjl0, T &gal0) {
    using O = FOps<T>;
    T z0 = T(0), phi0 = T(0), jl0 = jl0c, al0 = al0c;
    if (init) {
        z0 = O::fma(-L.nu0, p0.u, L.dnu) * inv_aa;
        phi0 = norm * O::ex(-(z0 * z0));
        jl0 = p0.jc * phi0;
        al0 = p0.dc * phi0;
    }
    const T jl1 = p1.jc * phi1, al1 = p1.dc * phi1;
    const T jh0 = p0.jd + jl0, jh1 = p1.jd + jl1;
    const T ah0 = p0.ad + al0, ah1 = p1.ad + al1;
    const T hd = T(0.5) * ds;
    const T dtau = O::fma(p1.sig, ah1, p0.sig * ah0) * hd;
    const T theomax = O::fma(p1.sig, jh1, p0.sig * jh0) * hd;
    T wa, wb, xp, ra;
    qdr_weights_f<T>(dtau, wa, wb, xp, ra);
    const T q1 = ah1 > T(0) ? O::sdiv(jh1, ah1) : T(0);
    const T q0 = ah0 > T(0) ? O::sdiv(jh0, ah0) : T(0);
    const T s1 = ah0 > T(0) ? q0 : (ah1 > T(0) ? q1 : T(0));
    const T s2 = ah1 > T(0) ? q1 : (ah0 > T(0) ? q0 : T(0));
    const bool thick = dtau > T(1e-9f);
    const T qraw = thick ? O::fma(wa, s1, wb * s2) : theomax;

    // I_out = I_in*xp + min(qraw, theomax)
    const T lam_in = lam * xp;
    const T g_xp = lam * I_in;
    T g_q = T(0), g_th = T(0);
    if (qraw < theomax) g_q = lam; else g_th = lam;
    T g_wa = T(0), g_wb = T(0), g_s1 = T(0), g_s2 = T(0);
    if (thick) {
        g_wa = g_q * s1; g_wb = g_q * s2;
        g_s1 = g_q * wa; g_s2 = g_q * wb;
    } else {
        g_th += g_q;
    }
    // weights: db/dx = a/x, da/dx = xp - a/x, dxp/dx = -xp (thin: 1/2, 1/2, -1)
    T g_tau;
    if (sizeof(T) == 4 || dtau > T(1.0e-6)) g_tau = -g_xp * xp + g_wa * (xp - ra) + g_wb * ra;
    else g_tau = -g_xp + T(0.5) * (g_wa + g_wb);
    // source-function fall-backs (transfer.F:1528-1541)
    T g_q0 = T(0), g_q1 = T(0);
    if (ah0 > T(0)) g_q0 += g_s1; else if (ah1 > T(0)) g_q1 += g_s1;
    if (ah1 > T(0)) g_q1 += g_s2; else if (ah0 > T(0)) g_q0 += g_s2;
    T g_jh0 = T(0), g_jh1 = T(0), g_ah0 = T(0), g_ah1 = T(0);
    // divided, not multiplied by 1/ah: ah reaches the denormal range in the
    // line wings of dust-free cells, where 1/ah overflows while g_q/ah does not
    if (ah1 > T(0) && g_q1 != T(0)) {
        const T g = O::sdiv(g_q1, ah1);
        g_jh1 += g;
        g_ah1 -= g * q1;
    }
    if (ah0 > T(0) && g_q0 != T(0)) {
        const T g = O::sdiv(g_q0, ah0);
        g_jh0 += g;
        g_ah0 -= g * q0;
    }
    // theomax and dtau
    const T gth_h = g_th * hd, gtau_h = g_tau * hd;
    g_jh0 += gth_h * p0.sig;
    g_jh1 += gth_h * p1.sig;
    g_ah0 += gtau_h * p0.sig;
    g_ah1 += gtau_h * p1.sig;
    g0.sig += gth_h * jh0 + gtau_h * ah0;
    g1.sig += gth_h * jh1 + gtau_h * ah1;
    g_ds += T(0.5) * (g_th * (p0.sig * jh0 + p1.sig * jh1) + g_tau * (p0.sig * ah0 + p1.sig * ah1));
    g0.jd += g_jh0; g0.ad += g_ah0;
    g1.jd += g_jh1; g1.ad += g_ah1;
    // end point line values (also carried into the next element)
    const T g_jl1 = g_jh1 + gjl1, g_al1 = g_ah1 + gal1;
    g1.jc += g_jl1 * phi1;
    g1.dc += g_al1 * phi1;
    const T t1 = (g_jl1 * p1.jc + g_al1 * p1.dc) * phi1;      // dL/dphi1 * phi1
    g_lnaa += t1 * (T(1) - T(2) * z1 * z1);
    g1.u += t1 * T(2) * z1 * L.nu0 * inv_aa;
    if (init) {
        g0.jc += g_jh0 * phi0;
        g0.dc += g_ah0 * phi0;
        const T t0 = (g_jh0 * p0.jc + g_ah0 * p0.dc) * phi0;
        g_lnaa += t0 * (T(1) - T(2) * z0 * z0);
        g0.u += t0 * T(2) * z0 * L.nu0 * inv_aa;
        gjl0 = T(0);
        gal0 = T(0);
    } else {
        gjl0 = g_jh0;
        gal0 = g_ah0;
    }
    return lam_in;
}

// largest v over the wave (all lanes active), in a scalar register
__device__ inline int wave_max(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

// profile argument and value of a point with a segment's constants, as
// elem_fast computes them
template <typename T>
__device__ inline void profile_at(const FastLineT<T> &L, T u, T inv_aa, T norm, T &z, T &phi) {
    z = FOps<T>::fma(-L.nu0, u, L.dnu) * inv_aa;
    phi = norm * FOps<T>::ex(-(z * z));
}

// dL/d(sub-point q at fraction e, with subpoint's normalised weights
// n0, n1 and 1/W) -> node gradients; returns dL/de
template <typename T>
__device__ inline T subpoint_rev(const FastRecT<T> &b, const FPtT<T> &p0, const FPtT<T> &p1,
                                 const FPtT<T> &q, const SubW<T> &w, const PGrad<T> &gq,
                                 PGrad<T> &g0, PGrad<T> &g1) {
    g0.jd += gq.jd * w.n0; g1.jd += gq.jd * w.n1;
    g0.ad += gq.ad * w.n0; g1.ad += gq.ad * w.n1;
    g0.jc += gq.jc * w.n0; g1.jc += gq.jc * w.n1;
    g0.dc += gq.dc * w.n0; g1.dc += gq.dc * w.n1;
    T g_e = gq.u * (p1.u - p0.u) + gq.sig * b.sigm * (b.rel1 - b.rel0);
    T acc = gq.jd * (b.rel1 * (p1.jd - q.jd) - b.rel0 * (p0.jd - q.jd));
    acc += gq.ad * (b.rel1 * (p1.ad - q.ad) - b.rel0 * (p0.ad - q.ad));
    acc += gq.jc * (b.rel1 * (p1.jc - q.jc) - b.rel0 * (p0.jc - q.jc));
    acc += gq.dc * (b.rel1 * (p1.dc - q.dc) - b.rel0 * (p0.dc - q.dc));
    // a point's u and sig do not enter the other quantities: d(n0, n1)/de
    // reach only jd, ad, jc, dc
    return g_e + acc * w.iw;
}

// Reverse of segment_fast for segment k (nodes k-1 -> k).  I_start is I
// before the segment (before a star blend), init_in the carry state
// entering it, jl0c/al0c the carried line values (valid when !init_in),
// z1/phi1 the profile at node k with this segment's constants.
// gjl/gal: in, dL/d(line values at node k carried into segment k+1);
// out, dL/d(carried values entering this segment).  isub is a stack of
// 32 entries.  Called by every lane of the wave (a lane with nothing to
// contribute passes lam = 0).  Returns dL/dI_start.
//
// A segment of m sub-points has m+1 elements; one that is integrated
// whole has m = 0.  Both run the same loop over the elements from the
// last, so that a wave whose lanes mix the two cases runs max(m)+1
// element reversals rather than that plus one (a lane joins at its own
// last element).  The end point of element
// e is the start point of element e+1, reversed just before: its
// position, interpolation weights and profile are carried over.
template <typename T>
__device__ inline T segment_rev(const FastRecT<T> &b, const FPtT<T> &p0, const FPtT<T> &p1,
                                const FastLineT<T> &L, T I_start, T lam, T istar, T starfract,
                                int init_in, T jl0c, T al0c, T z1, T phi1, T &gjl, T &gal,
                                PGrad<T> &g0, PGrad<T> &g1, T &g_lnaa, T &g_dls, T *isub) {
    T ds = b.ds;
    int init = init_in;
    const int fl = (int)b.flag;
    if (fl & NF_RESET) { ds = T(0); init = 1; }
    T I = I_start;
    if (fl & NF_BLEND) { I = (T(1) - starfract) * I + starfract * istar; init = 1; }
    SubGrid<T> sg{T(0), T(0), 1, 0};
    int m = 0, i1 = 1;              // sub-points i1 .. i1+m-1 (0 < s_i < ds)
    T inv_ds = T(0);
    if (subgrid_of<T>(b, ds, p0.u, L.uch, sg)) {
        inv_ds = FOps<T>::fdiv(T(1), ds);
        // s_i = sleft + (i-1) lg does not decrease with i, so the indices
        // with 0 < s_i < ds are one run inside [ilo, ihi], whose bounds are
        // within a step or two of it
        auto inside = [&](int i) {
            const T s = sg.sleft + (T)(i - 1) * sg.lg;
            return s > T(0) && s < ds;
        };
        int ia = sg.ilo, ib = sg.ihi;
        while (ia <= ib && !inside(ia)) ia++;
        while (ib >= ia && !inside(ib)) ib--;
        i1 = ia;
        m = ib >= ia ? ib - ia + 1 : 0;
    }
    // the wave steps through element indices together, so that the stack
    // of element start values is indexed by a wave-uniform e (registers)
    const int M = __any(m > 0) ? wave_max(m) : 0;
    if (M == 0) {
        // no lane of the wave sub-grids (most segments): one element each
        T g_h = T(0), gjl0, gal0;
        lam = elem_rev<T>(L, I, lam, ds, p0, p1, z1, phi1, b.inv_aa, b.norm, init, jl0c, al0c,
                          gjl, gal, g0, g1, g_h, g_lnaa, gjl0, gal0);
        if (fl & NF_BLEND) lam = lam * (T(1) - starfract);
        gjl = init ? T(0) : gjl0;
        gal = init ? T(0) : gal0;
        return lam;
    }
    {
        // forward over the sub-steps, keeping I at each element start
        FastCarryT<T> c;
        c.init = init;
        if (!init) carry_start<T>(p0, jl0c, al0c, c);
        int ovs = 0;
        FPtT<T> pp = p0;
        T sa = T(0);
        for (int e = 0; e < M; e++) {
            if (e < m) {
                const T s = sg.sleft + (T)(i1 + e - 1) * sg.lg;
                const FPtT<T> q = subpoint<T>(b, p0, p1, s * inv_ds);
                isub[e] = I;
                elem_fast<T>(L, I, s - sa, pp, q, b.inv_aa, b.norm, c, ovs);
                pp = q;
                sa = s;
            }
        }
    }
    // I: start of the last element.  Points: 0 = node k-1, j = sub-point
    // i1+j-1 (j = 1..m), m+1 = node k; element e runs from point e to e+1.
    // The sub-points of a lane are consecutive indices: s_i grows with i,
    // so those inside (0, ds) form one run.
    T gjl1 = gjl, gal1 = gal;       // dL/d(line values at the end point)
    PGrad<T> gB;                    // gradient of the end point
    pg_zero(gB);
    T gsB = T(0);                   // dL/ds of the end point (sub-points only)
    FPtT<T> B = p1;
    SubW<T> wB{T(1), T(0), T(1), T(1)};
    T zB = z1, phiB = phi1, sB = ds;
    for (int e = M; e >= 0; e--) {
        if (e > m) continue;        // this lane's last element comes later
        FPtT<T> A = p0;
        SubW<T> wA{T(0), T(1), T(0), T(1)};
        T zA = T(0), phiA = T(0), jlA = jl0c, alA = al0c, sA = T(0);
        if (e > 0) {
            sA = sg.sleft + (T)(i1 + e - 2) * sg.lg;
            A = subpoint_w<T>(b, p0, p1, sA * inv_ds, wA);
            profile_at<T>(L, A.u, b.inv_aa, b.norm, zA, phiA);
            jlA = A.jc * phiA;
            alA = A.dc * phiA;
        }
        PGrad<T> gA;
        pg_zero(gA);
        T g_h = T(0), gjl0, gal0;
        T Iin = I;
        if (e < m) Iin = isub[e];
        lam = elem_rev<T>(L, Iin, lam, sB - sA, A, B, zB, phiB, b.inv_aa,
                          b.norm, e == 0 ? init : 0, jlA, alA, gjl1, gal1, gA, gB, g_h, g_lnaa,
                          gjl0, gal0);
        if (e < m) {
            // end point B is sub-point i1+e: to the nodes, and its position
            // s = sleft + (i-1) lg = s_c + dls (0.2 (i-1) - 3) to the width
            gsB += g_h;
            gsB += subpoint_rev<T>(b, p0, p1, B, wB, gB, g0, g1) * inv_ds;
            g_dls += gsB * (T(-3) + (T)(i1 + e - 1) * T(0.2));
        } else {
            g1.jd += gB.jd; g1.ad += gB.ad; g1.jc += gB.jc; g1.dc += gB.dc;
            g1.u += gB.u; g1.sig += gB.sig;
        }
        // the start point becomes the next end point
        gB = gA;
        gsB = -g_h;
        B = A;
        wB = wA;
        zB = zA;
        phiB = phiA;
        sB = sA;
        gjl1 = gjl0;
        gal1 = gal0;
    }
    // point 0 is node k-1
    g0.jd += gB.jd; g0.ad += gB.ad; g0.jc += gB.jc; g0.dc += gB.dc;
    g0.u += gB.u; g0.sig += gB.sig;
    if (fl & NF_BLEND) lam = lam * (T(1) - starfract);
    gjl = init ? T(0) : gjl1;
    gal = init ? T(0) : gal1;
    return lam;
}

}  // namespace rl
