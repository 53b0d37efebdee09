// Single-precision forward trace on split records: the batched spectrum
// path of the FP32 mode (charintline, telescope.F:3889-4312, with
// clever_integrate_element_linedust line.F:4636-4848,
// integrate_element_linedust line.F:4515-4624, voigt_profile
// line.F:2280-2314 and qdr_src_2 transfer.F:1498-1571).
//
// What a segment needs is split by what it depends on:
//
//   SegHot / SegAux  per trajectory node, independent of the line: built
//                    once per inclination (path length, projected
//                    velocity, the line width's profile scale and
//                    normalisation, the sub-grid length, node flags);
//   NodeLine         per item and node: the node's dust and line
//                    coefficients, scaled (16 bytes).
//
// With e = exp(-z^2) the unnormalised Gaussian profile, the profile of a
// segment with mean width lwav at channel velocity uch (dnu/nu0) is
// phi = (ng/|nu0|) e with z = (uch - u) / (3.33567e-6 lwav), ng =
// 0.56419583546 / (3.33567e-6 lwav): the same quantity voigt_profile forms
// from aa = 3.33567e-6 |nu0| lwav, its 1/|nu0| moved into the line
// coefficients.  The argument is carried pre-multiplied by sqrt(log2 e)
// (ag) so that e is one hardware exp2.
//
// The hot loop reads 32 bytes per segment (SegHot + NodeLine); a segment
// that may be sub-gridded, carries a node flag, or starts the ray takes
// the general path, which also reads SegAux and the start node's record.
#pragma once
#include "rl_device.h"

namespace rl {

struct alignas(16) SegHot {
    float hds;   // ds/2 (0 on a hole chord); sign bit set: the general path
    float ag;    // sqrt(log2 e) / (3.33567e-6 lwav)  [per unit of u]
    float uag;   // u_k * ag
    float ng;    // 0.56419583546 / (3.33567e-6 lwav)
};

// (ds is SegHot's: 2 |hds| is the single-precision ds exactly, the
// halving and doubling being exact in binary)
struct alignas(16) SegAux {
    float u;       // Omega.v/c at node k
    float dls;     // sub-grid line length ds/x (line.F:4733-4738), < 0: never
    float inv_du;  // 1 / (u_k - u_{k-1})
    int flag;      // NF_* bits of node k
};

// one 16-byte load (read field by field, the record took four)
__device__ inline SegAux load_aux(const SegAux *p) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    SegAux a;
    a.u = v.x;
    a.dls = v.y;
    a.inv_du = v.z;
    a.flag = __float_as_int(v.w);
    return a;
}

// j = |sig| (jdh + jcl ng e), alpha = |sig| (adh + dcl ng e); sig is the
// node's scale (its dust opacity where there is dust, adh = 1); the sign
// bit of sig is set where there is no dust (adh = 0).  jcl and dcl are the
// line coefficients of line.F:4578-4588 over sig |nu0|.
struct alignas(16) NodeLine {
    float sig, jdh, jcl, dcl;
};

constexpr float SQRT_LOG2E = 1.20112240878644977f;

__device__ inline float nl_adh(float sig) { return __float_as_int(sig) < 0 ? 0.0f : 1.0f; }
__device__ inline bool seg_general(float hds) { return __float_as_int(hds) < 0; }
__device__ inline float ex2(float x) { return __builtin_amdgcn_exp2f(x); }

// the values an element needs of its start point, carried from the element
// that ended there: sig*alpha/sig and sig*j/sig (the products the optical
// depth and the bracket add), alpha/sig (its sign picks the source
// functions) and the source function
struct Car32 {
    float sa0 = 0.0f, sj0 = 0.0f, ah0 = 0.0f, S0 = 0.0f;
};

// a point of the general path: scaled values (adh explicit)
struct Pt32 {
    float u, sig, jdh, adh, jcl, dcl;
};

__device__ inline Pt32 pt_of(float u, const NodeLine &q) {
    return Pt32{u, fabsf(q.sig), q.jdh, nl_adh(q.sig), q.jcl, q.dcl};
}

// the carry of a start point p under a segment's profile (ag, ng)
__device__ inline void carry32(const Pt32 &p, float uch, float ag, float ng, Car32 &c) {
    const float z = (uch - p.u) * ag;
    const float e = ex2(-(z * z));
    const float ah = fmaf(p.dcl * ng, e, p.adh);
    const float jh = fmaf(p.jcl * ng, e, p.jdh);
    c.sa0 = p.sig * ah;
    c.sj0 = p.sig * jh;
    c.ah0 = ah;
    c.S0 = ah > 0.0f ? FOps<float>::sdiv(jh, ah) : 0.0f;
}

// One element from the carried start point to an end point with scaled
// values (sg, adh, jdh) and profile-weighted line coefficients jcn = jcl
// ng, dcn = dcl ng at e: I <- I exp(-dtau) + min(Q, J_max) (qdr_src_2,
// transfer.F:1498-1571); ovf is set on an overshoot (Q > 2 J_max).
__device__ inline float elem32(float I, float hds, float e, float jcn, float dcn, float sg,
                               float adh, float jdh, Car32 &c, bool &ovf) {
    const float ah1 = fmaf(dcn, e, adh);
    const float jh1 = fmaf(jcn, e, jdh);
    const float sa1 = sg * ah1, sj1 = sg * jh1;
    const float dtau = (c.sa0 + sa1) * hds;
    const float theomax = (c.sj0 + sj1) * hds;
    float wa, wb, xp, ra;
    qdr_weights_f<float>(dtau, wa, wb, xp, ra);
    // source functions with the fall-backs of transfer.F:1528-1541
    const float q1 = ah1 > 0.0f ? FOps<float>::sdiv(jh1, ah1) : 0.0f;
    const float q0 = c.S0;
    const float s1 = c.ah0 > 0.0f ? q0 : (ah1 > 0.0f ? q1 : 0.0f);
    const float s2 = ah1 > 0.0f ? q1 : (c.ah0 > 0.0f ? q0 : 0.0f);
    float q = dtau > 1e-9f ? fmaf(wa, s1, wb * s2) : theomax;   // 1e-9 is REAL*4
    ovf = ovf || (q > 2.0f * theomax);
    q = fminf(q, theomax);
    c.sa0 = sa1;
    c.sj0 = sj1;
    c.ah0 = ah1;
    c.S0 = q1;
    return fmaf(I, xp, q);
}

// an element ending at point p of the general path
__device__ inline float elem_pt32(float I, float hds, const Pt32 &p, float uch, float ag,
                                  float ng, Car32 &c, bool &ovf, float &e_out) {
    const float z = (uch - p.u) * ag;
    const float e = ex2(-(z * z));
    e_out = e;
    return elem32(I, hds, e, p.jcl * ng, p.dcl * ng, p.sig, p.adh, p.jdh, c, ovf);
}

// The general path of segment k (nodes k-1 -> k): node flags (hole chord,
// central-beam star blend, star intensity), the start point when the
// element starts afresh (init), and the sub-grid of
// clever_integrate_element_linedust (line.F:4706-4800): points s_i =
// sleft + (i-1) lg, i = 1..31, used where 0 < s_i < ds, the quantities
// interpolated linearly in s, the width held at the segment's mean.
__device__ inline float seg_general32(float I, const SegHot &h, const SegAux &a0,
                                      const SegAux &a1, const NodeLine &q0, const NodeLine &q1,
                                      float uch, float istar, float sf, int &init, Car32 &c,
                                      bool &ovf, int &maser) {
    float ds = 2.0f * fabsf(h.hds);                     // (float)ds, hds = (float)(ds/2)
    const int fl = a1.flag;
    if (fl & NF_RESET) { ds = 0.0f; init = 1; }
    if (fl & NF_BLEND) { I = (1.0f - sf) * I + sf * istar; init = 1; }
    if (fl & NF_STAR) { I = istar; init = 1; }
    const Pt32 p0 = pt_of(a0.u, q0), p1 = pt_of(a1.u, q1);
    if (init) carry32(p0, uch, h.ag, h.ng, c);
    init = 0;
    float sa = 0.0f, e1 = 0.0f;
    if (a1.dls >= 0.0f) {                                 // 6x > 1 (line.F:4715)
        const float s_c = ds * (uch - p0.u) * a1.inv_du;
        const float sright = s_c + 3.0f * a1.dls, sleft = s_c - 3.0f * a1.dls;
        if (sright > 0.0f && sleft < ds) {
            const float lg = (sright - sleft) * (1.0f / 30.0f);
            const float f0 = FOps<float>::fdiv(-sleft, lg), f1 = FOps<float>::fdiv(ds - sleft, lg);
            const int ilo = f0 > 0.0f ? max(1, (int)fminf(f0, 40.0f)) : 1;
            const int ihi = f1 < 40.0f ? min(31, (int)fmaxf(f1, -2.0f) + 2) : 31;
            // the nodes' scales relative to the larger; where both are below
            // the single-precision range (no optical depth to speak of) the
            // points are weighted equally
            const float sm = fmaxf(p0.sig, p1.sig);
            const float rel0 = sm > 0.0f ? p0.sig / sm : 1.0f;
            const float rel1 = sm > 0.0f ? p1.sig / sm : 1.0f;
            const float inv_ds = FOps<float>::fdiv(1.0f, ds);
            for (int i = ilo; i <= ihi; i++) {
                const float s = sleft + (float)(i - 1) * lg;
                if (s > 0.0f && s < ds) {
                    // linear interpolation of the unscaled quantities
                    // (line.F:4775-4790) through relative weights
                    const float e = s * inv_ds, ep = 1.0f - e;
                    float w0 = ep * rel0, w1 = e * rel1;
                    const float W = w0 + w1, iw = FOps<float>::fdiv(1.0f, W);
                    w0 *= iw;
                    w1 *= iw;
                    const Pt32 b{ep * p0.u + e * p1.u, sm * W, w0 * p0.jdh + w1 * p1.jdh,
                                 w0 * p0.adh + w1 * p1.adh, w0 * p0.jcl + w1 * p1.jcl,
                                 w0 * p0.dcl + w1 * p1.dcl};
                    float eb;
                    I = elem_pt32(I, 0.5f * (s - sa), b, uch, h.ag, h.ng, c, ovf, eb);
                    sa = s;
                }
            }
        }
    }
    I = elem_pt32(I, 0.5f * (ds - sa), p1, uch, h.ag, h.ng, c, ovf, e1);
    // maser flag (telescope.F:4295-4297): line alpha * ds < -0.01 at the end
    if (p1.sig * (p1.dcl * h.ng * e1) * ds < -0.01f) maser = 1;
    return I;
}

// charintline on split records: one channel (uch = dnu/nu0) of one ray of n
// nodes from the outer boundary intensity I0
__device__ inline float trace_split(const SegHot *__restrict__ hot, const SegAux *__restrict__ aux,
                                    const NodeLine *__restrict__ line, int n, float uch, float I0,
                                    float istar, float sf, int &ovs, int &maser) {
    float I = I0;
    if (n < 2) return I;
    Car32 c;
    int init = 1;
    bool ovf = false;
    // segment 1 starts the ray: the general path reads its start point
    I = seg_general32(I, hot[1], aux[0], load_aux(aux + 1), line[0], line[1], uch, istar, sf,
                      init, c, ovf, maser);
    auto segment = [&](int k, const SegHot &h, const NodeLine &q) {
        if (seg_general(h.hds)) {
            I = seg_general32(I, h, aux[k - 1], load_aux(aux + k), line[k - 1], q, uch, istar, sf,
                              init, c, ovf, maser);
        } else {
            const float z = fmaf(uch, h.ag, -h.uag);
            const float e = ex2(-(z * z));
            const float dcn = q.dcl * h.ng;
            I = elem32(I, h.hds, e, q.jcl * h.ng, dcn, fabsf(q.sig), nl_adh(q.sig), q.jdh, c, ovf);
            // maser flag; the line coefficient is negative only where the
            // populations are inverted
            if (q.dcl < 0.0f && fabsf(q.sig) * (dcn * e) * (2.0f * h.hds) < -0.01f) maser = 1;
        }
    };
    // Segments 2.. in pairs, the records in two register sets (A: segments
    // k, k+2, ..; B: k+1, k+3, ..): a set is reloaded right after its
    // segment is done, so the load overlaps the other set's segment and no
    // copies rotate the registers.
    if (n > 2) {
        int k = 2;
        SegHot hA = hot[k];
        NodeLine qA = line[k];
        const int k1 = k + 1 < n ? k + 1 : k;
        SegHot hB = hot[k1];
        NodeLine qB = line[k1];
        for (;;) {
            segment(k, hA, qA);
            if (k + 1 >= n) break;
            const int ka = k + 2 < n ? k + 2 : n - 1;
            hA = hot[ka];
            qA = line[ka];
            segment(k + 1, hB, qB);
            if (k + 2 >= n) break;
            const int kb = k + 3 < n ? k + 3 : n - 1;
            hB = hot[kb];
            qB = line[kb];
            k += 2;
        }
    }
    ovs += ovf ? 1 : 0;
    return I;
}

}  // namespace rl
