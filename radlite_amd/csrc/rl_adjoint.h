// Reverse-mode (adjoint) formal solution on scaled records.
//
// The forward model per (ray, channel) is the recurrence of charintline
// (telescope.F:3889-4312): I_out = I_in * xp + min(Q, theomax) per
// element (qdr_src_2, transfer.F:1498-1571), elements being whole
// segments or the sub-grid steps of clever_integrate_element_linedust
// (line.F:4636-4848).  Its adjoint runs the segments backwards with
// lambda = dL/dI, recomputing I at each segment start from checkpoints
// taken every ADJ_K segments by the forward pass, and I at each sub-step
// start by recomputing the segment.  Everything that depends on the
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
// ds, end points, profile constants, init and the carried line values
// jl0c/al0c used when init == 0), lam = dL/dI_out, and gjl1/gal1 =
// dL/d(line j, alpha at the end point) arriving from the next element,
// which carried them.  Accumulates into g0/g1 (the two points), g_ds and
// g_lnaa (= dL/d ln inv_aa); returns dL/dI_in and, when init == 0, the
// gradient on the carried start values in gjl0/gal0.
template <typename T>
__device__ inline T elem_rev(const FastLineT<T> &L, T I_in, T lam, T ds, const FPtT<T> &p0,
                             const FPtT<T> &p1, T inv_aa, T norm, int init, T jl0c, T al0c,
                             T gjl1, T gal1, PGrad<T> &g0, PGrad<T> &g1, T &g_ds, T &g_lnaa,
                             T &gjl0, T &gal0) {
    T z0 = T(0), phi0 = T(0), jl0 = jl0c, al0 = al0c;
    if (init) {
        z0 = (L.dnu - L.nu0 * p0.u) * inv_aa;
        phi0 = norm * FOps<T>::ex(-(z0 * z0));
        jl0 = p0.jc * phi0;
        al0 = p0.dc * phi0;
    }
    const T z1 = (L.dnu - L.nu0 * p1.u) * inv_aa;
    const T phi1 = norm * FOps<T>::ex(-(z1 * z1));
    const T jl1 = p1.jc * phi1, al1 = p1.dc * phi1;
    const T jh0 = p0.jd + jl0, jh1 = p1.jd + jl1;
    const T ah0 = p0.ad + al0, ah1 = p1.ad + al1;
    const T a0 = p0.sig * ah0, a1 = p1.sig * ah1;
    const T dtau = T(0.5) * (a0 + a1) * ds;
    const T theomax = T(0.5) * (p0.sig * jh0 + p1.sig * jh1) * ds;
    T wa, wb, xp, ra;
    qdr_weights_f<T>(dtau, wa, wb, xp, ra);
    const T q1 = ah1 > T(0) ? FOps<T>::sdiv(jh1, ah1) : T(0);
    const T q0 = ah0 > T(0) ? FOps<T>::sdiv(jh0, ah0) : T(0);
    const T s1 = ah0 > T(0) ? q0 : (ah1 > T(0) ? q1 : T(0));
    const T s2 = ah1 > T(0) ? q1 : (ah0 > T(0) ? q0 : T(0));
    const bool thick = dtau > T(1e-9f);
    const T qraw = thick ? wa * s1 + wb * s2 : theomax;

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
        const T g = FOps<T>::sdiv(g_q1, ah1);
        g_jh1 += g;
        g_ah1 -= g * q1;
    }
    if (ah0 > T(0) && g_q0 != T(0)) {
        const T g = FOps<T>::sdiv(g_q0, ah0);
        g_jh0 += g;
        g_ah0 -= g * q0;
    }
    // theomax and dtau
    const T hd = T(0.5) * ds;
    g_jh0 += g_th * hd * p0.sig;
    g_jh1 += g_th * hd * p1.sig;
    g_ah0 += g_tau * hd * p0.sig;
    g_ah1 += g_tau * hd * p1.sig;
    g0.sig += hd * (g_th * jh0 + g_tau * ah0);
    g1.sig += hd * (g_th * jh1 + g_tau * ah1);
    g_ds += T(0.5) * (g_th * (p0.sig * jh0 + p1.sig * jh1) + g_tau * (a0 + a1));
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

// dL/d(sub-point at fraction e) -> node gradients and dL/de
template <typename T>
__device__ inline T subpoint_rev(const FastRecT<T> &b, const FPtT<T> &p0, const FPtT<T> &p1,
                                 T e, const PGrad<T> &gq, PGrad<T> &g0, PGrad<T> &g1) {
    const T ep = T(1) - e;
    const T w0 = ep * b.rel0, w1 = e * b.rel1, W = w0 + w1;
    const T iw = FOps<T>::fdiv(T(1), W);
    const T n0 = w0 * iw, n1 = w1 * iw;
    const T xq[4] = {n0 * p0.jd + n1 * p1.jd, n0 * p0.ad + n1 * p1.ad,
                     n0 * p0.jc + n1 * p1.jc, n0 * p0.dc + n1 * p1.dc};
    const T x0[4] = {p0.jd, p0.ad, p0.jc, p0.dc};
    const T x1[4] = {p1.jd, p1.ad, p1.jc, p1.dc};
    const T gx[4] = {gq.jd, gq.ad, gq.jc, gq.dc};
    g0.jd += gx[0] * n0; g1.jd += gx[0] * n1;
    g0.ad += gx[1] * n0; g1.ad += gx[1] * n1;
    g0.jc += gx[2] * n0; g1.jc += gx[2] * n1;
    g0.dc += gx[3] * n0; g1.dc += gx[3] * n1;
    T g_e = gq.u * (p1.u - p0.u) + gq.sig * b.sigm * (b.rel1 - b.rel0);
    for (int f = 0; f < 4; f++)
        g_e += gx[f] * (b.rel1 * (x1[f] - xq[f]) - b.rel0 * (x0[f] - xq[f])) * iw;
    return g_e;
}

// Reverse of segment_fast for segment k (nodes k-1 -> k).  I_start is I
// before the segment (before a star blend), init_in the carry state
// entering it, jl0c/al0c the carried line values (valid when !init_in).
// gjl/gal: in, dL/d(line values at node k carried into segment k+1);
// out, dL/d(carried values entering this segment).  isub is a scratch
// stack of at least 33 entries.  Returns dL/dI_start.
template <typename T>
__device__ inline T segment_rev(const FastRecT<T> &b, const FPtT<T> &p0, const FPtT<T> &p1,
                                const FastLineT<T> &L, T I_start, T lam, T istar, T starfract,
                                int init_in, T jl0c, T al0c, T &gjl, T &gal,
                                PGrad<T> &g0, PGrad<T> &g1, T &g_lnaa, T &g_dls, T *isub) {
    T ds = b.ds;
    int init = init_in;
    const int fl = (int)b.flag;
    if (fl & NF_RESET) { ds = T(0); init = 1; }
    T I = I_start;
    if (fl & NF_BLEND) { I = (T(1) - starfract) * I + starfract * istar; init = 1; }
    T g_ds = T(0);     // segment length is geometry: not a parameter
    T gjl_c, gal_c;
    SubGrid<T> sg;
    if (subgrid_of<T>(b, ds, p0.u, L.uch, sg)) {
        // forward over the sub-steps, keeping I at each element start
        const T inv_ds = FOps<T>::fdiv(T(1), ds);
        int m = 0;
        {
            FastCarryT<T> c{jl0c, al0c, T(0), init};
            if (!init) {
                const T ah = p0.ad + al0c;
                c.S0 = ah > T(0) ? FOps<T>::sdiv(p0.jd + jl0c, ah) : T(0);
            }
            int ovs = 0;
            FPtT<T> pp = p0;
            T sa = T(0);
            for (int i = sg.ilo; i <= sg.ihi; i++) {
                const T s = sg.sleft + (T)(i - 1) * sg.lg;
                if (s > T(0) && s < ds) {
                    const FPtT<T> q = subpoint<T>(b, p0, p1, s * inv_ds);
                    isub[m++] = I;
                    elem_fast<T>(L, I, s - sa, pp, q, b.inv_aa, b.norm, c, ovs);
                    pp = q;
                    sa = s;
                }
            }
            isub[m] = I;     // start of the last element
        }
        // backwards over the elements: element e runs from point e to e+1
        // (point 0 = node k-1, point m+1 = node k)
        T gjl1 = gjl, gal1 = gal;        // on the end point of the current element
        PGrad<T> gB;                      // accumulated gradient of the current end point
        pg_zero(gB);
        T gsB = T(0);                     // dL/ds of that end point (sub-points only)
        int iB = sg.ihi + 1;              // index of the end point's sub-grid node
        T sB = ds;
        for (int e = m; e >= 0; e--) {
            // start point of element e
            int iA = -1;
            T sA = T(0);
            if (e > 0) {
                for (int i = iB - 1; i >= sg.ilo; i--) {
                    const T s = sg.sleft + (T)(i - 1) * sg.lg;
                    if (s > T(0) && s < ds) { iA = i; sA = s; break; }
                }
            }
            const FPtT<T> A = e > 0 ? subpoint<T>(b, p0, p1, sA * inv_ds) : p0;
            const FPtT<T> B = e < m ? subpoint<T>(b, p0, p1, sB * inv_ds) : p1;
            int einit = e == 0 ? init : 0;
            T jlA = jl0c, alA = al0c;
            if (e > 0) {
                const T z = (L.dnu - L.nu0 * A.u) * b.inv_aa;
                const T phi = b.norm * FOps<T>::ex(-(z * z));
                jlA = A.jc * phi;
                alA = A.dc * phi;
            }
            PGrad<T> gA;
            pg_zero(gA);
            T g_h = T(0), gjl0, gal0;
            lam = elem_rev<T>(L, isub[e], lam, sB - sA, A, B, b.inv_aa, b.norm, einit, jlA, alA,
                              gjl1, gal1, gA, gB, g_h, g_lnaa, gjl0, gal0);
            // finish the end point
            if (e < m) {
                gsB += g_h;
                const T g_e = subpoint_rev<T>(b, p0, p1, sB * inv_ds, gB, g0, g1);
                gsB += g_e * inv_ds;
                g_dls += gsB * (T(-3) + (T)(iB - 1) * T(0.2));
            } else {
                g1.jd += gB.jd; g1.ad += gB.ad; g1.jc += gB.jc; g1.dc += gB.dc;
                g1.u += gB.u; g1.sig += gB.sig;
                g_ds += g_h;
            }
            // the start point becomes the next end point
            gB = gA;
            gsB = -g_h;
            iB = iA;
            sB = sA;
            gjl1 = gjl0;
            gal1 = gal0;
        }
        // point 0 is node k-1
        g0.jd += gB.jd; g0.ad += gB.ad; g0.jc += gB.jc; g0.dc += gB.dc;
        g0.u += gB.u; g0.sig += gB.sig;
        gjl_c = gjl1;
        gal_c = gal1;
    } else {
        lam = elem_rev<T>(L, I, lam, ds, p0, p1, b.inv_aa, b.norm, init, jl0c, al0c, gjl, gal,
                          g0, g1, g_ds, g_lnaa, gjl_c, gal_c);
    }
    (void)g_ds;
    if (fl & NF_BLEND) lam = lam * (T(1) - starfract);
    if (init) { gjl_c = T(0); gal_c = T(0); }
    gjl = gjl_c;
    gal = gal_c;
    return lam;
}

}  // namespace rl
