// Ray/cell crossing geometry on the device, FP64, one thread per ray.
//
// Restates make_trajectory_c (RADLITE/telescope.F:2787-3720) with the
// reference's root formulas, eps filters and extra-point rule, but
// without its per-ray arrays: the theta-cone and R-sphere crossing
// lists are generated on demand from their index, the extra points live
// in a small private list, and the three-way merge streams nodes straight
// to the caller's sink.  The cones a ray crosses are those with
// tan^2(Theta) above a threshold (the discriminant is
// 4 (tan^2 (cos^2 th0 x0^2 + sin^2 th0 z0^2) - sin^2 th0 x0^2)), a
// contiguous range of cone indices; their roots are then recomputed from
// the index when the merge reaches them, with the same arithmetic.  Only
// when rounding breaks that contiguity are the roots kept, in a strided
// scratch buffer.
#pragma once
#include "rl_device.h"

namespace rl {

struct GridView {
    int nr, ny;               // ny = mirrored theta count
    const double *rc;         // rc[i+1] = R_i, i=-1..nr+2
    const double *tc;         // tc[i+1] = Theta_i, i=-1..ny+2
    double rstar;
    int t_sorted;             // Theta_1..Theta_ny non-decreasing
    __device__ double R(int i) const { return rc[i + 1]; }
    __device__ double TH(int i) const { return tc[i + 1]; }
};

// A crossing as the merge emits it: which surface was crossed and where.
// icross 1: R = R_idx sphere; 2: Theta = Theta_idx cone; 3: extra point
// (idx = its radial cell when it is the R-extremum, rext = 1).
struct Crossing {
    double s;
    int icross, idx, rext;
};

struct Traj {
    const GridView &g;
    const double *tan2;       // tan^2(Theta_iy), iy = 1..ny/2, from the host libm
    double x0, z0, theta0;
    double costh0, sinth0, costh02, sinth02, pitheta0;
    int iyeq, iyend;
    // theta cones
    double *sar1, *sar2;
    int *iyar;
    int stride;
    int iy0 = 0;              // > 0: cone is has index iy0 + is - 1 (roots recomputed)
    int isnr, isdblnr, iup, branchA, ith_amount;
    // R spheres
    int ir_min, ir_amount, half;
    double bimpact;

    __device__ Traj(const GridView &g_, const double *tan2_, double x0_, double z0_,
                    double th0, double cth0, double sth0, double *s1, double *s2, int *iy,
                    int strd)
        : g(g_), tan2(tan2_), x0(x0_), z0(z0_), theta0(th0), costh0(cth0), sinth0(sth0),
          sar1(s1), sar2(s2), iyar(iy), stride(strd) {}

    __device__ double thr(double s) const {
        return sqrt(x0 * x0 + z0 * z0 + s * s + 2.0 * z0 * costh0 * s);
    }
    __device__ double theta_of_s(double s) const {
        double th = atan(sqrt(x0 * x0 + sinth02 * s * s) / (z0 + costh0 * s));
        if (th < 0.0) th = th + PICONST;
        return th;
    }
    __device__ int hunt_r(double x) const {
        return hunt_bisect([&](int i) { return g.R(i); }, g.nr, x);
    }
    __device__ int hunt_t(double x) const {
        return hunt_bisect([&](int i) { return g.TH(i); }, g.ny, x);
    }
    // roots of cone iy, ordered (telescope.F:2968-3000); sd > 0 is known
    __device__ void cone_roots(int iy, double &s1, double &s2) const {
        double tanth2 = tan2[iy];
        double a = tanth2 * costh02 - sinth02;
        double b = 2.0 * tanth2 * costh0 * z0;
        double c = tanth2 * z0 * z0 - x0 * x0;
        double sd = sqrt(b * b - 4.0 * a * c);
        s1 = (-b - sd) / (2.0 * a);
        s2 = (-b + sd) / (2.0 * a);
        if (s1 > s2) { double d = s1; s1 = s2; s2 = d; }
    }
    // one root of cone iy: the smaller (lo) or the larger; the pair is
    // ordered by the sign of the leading coefficient, so each is the same
    // expression cone_roots evaluates
    __device__ double cone_root(int iy, bool lo) const {
        double tanth2 = tan2[iy];
        double a = tanth2 * costh02 - sinth02;
        double b = 2.0 * tanth2 * costh0 * z0;
        double c = tanth2 * z0 * z0 - x0 * x0;
        double sd = sqrt(b * b - 4.0 * a * c);
        return (lo == (a > 0.0)) ? (-b - sd) / (2.0 * a) : (-b + sd) / (2.0 * a);
    }
    __device__ double S1(int is) const {
        if (iy0 > 0) return cone_root(iy0 + is - 1, true);
        return sar1[(size_t)is * stride];
    }
    __device__ double S2(int is) const {
        if (iy0 > 0) return cone_root(iy0 + is - 1, false);
        return sar2[(size_t)is * stride];
    }
    __device__ int IY(int is) const { return iy0 > 0 ? iy0 + is - 1 : iyar[(size_t)is * stride]; }

    // th list entry iss (1 <= iss <= ith_amount): its cone is, root (1: the
    // smaller) and cone number itheta, ordering of telescope.F:3024-3194
    __device__ void th_index(int iss, int &is, int &root, int &itheta) const {
        int upq;
        const int n = isnr, d = isdblnr;
        if (branchA) {
            if (iss <= n - d) { is = d + iss; root = 1; upq = 0; }
            else if (iss <= 2 * n - 2 * d) { is = n - (iss - (n - d + 1)); root = 2; upq = 1; }
            else if (iss <= 2 * n - d) { is = d - (iss - (2 * n + 1 - 2 * d)); root = 1; upq = 1; }
            else { is = 1 + (iss - (2 * n + 1 - d)); root = 2; upq = 1; }
        } else {
            if (iss <= d) { is = d - (iss - 1); root = 1; upq = 0; }
            else if (iss <= 2 * d) { is = 1 + (iss - d - 1); root = 2; upq = 0; }
            else if (iss <= n + d) { is = d + 1 + (iss - 2 * d - 1); root = 1; upq = 0; }
            else { is = n - (iss - n - d - 1); root = 2; upq = 1; }
        }
        const int iy = IY(is);
        if (upq) itheta = iup == 1 ? iy : iyend + 1 - iy;
        else itheta = iup == 1 ? iyend + 1 - iy : iy;
    }
    // th list entry iss (1-based)
    __device__ void th_entry(int iss, double &s, int &itheta) const {
        if (iss > ith_amount) { s = 1.0e30; itheta = 0; return; }
        int is, root;
        th_index(iss, is, root, itheta);
        s = root == 1 ? S1(is) : S2(is);
    }
    __device__ double th_s(int iss) const {
        double s; int it;
        th_entry(iss, s, it);
        return s;
    }
    // r list entry k (1-based), telescope.F:3284-3335
    __device__ void r_entry(int k, double &s, double &r, int &ix) const {
        if (k > ir_amount) { s = 1.0e30; r = 0.0; ix = 0; return; }
        if (k < 1) { s = 0.0; r = 0.0; ix = 0; return; }   // r_s(0): see header
        int is = k <= half ? k : ir_amount + 1 - k;
        ix = g.nr - is + 1;
        r = ix == 0 ? g.rstar : g.R(ix);
        double a = 1.0, b = 2.0 * costh0 * z0, c = x0 * x0 + z0 * z0 - r * r;
        double sd = sqrt(b * b - 4.0 * a * c + TELESC_EPS * b * b);
        s = k <= half ? (-b - sd) / (2.0 * a) : (-b + sd) / (2.0 * a);
    }
    __device__ double r_s(int k) const {
        double s, r; int ix;
        r_entry(k, s, r, ix);
        return s;
    }
    // The merge's next entry: th_entry(k) into (ths, thit) when th, else
    // r_entry(k) (k >= 1) into (rs, rr, rix), with the same arithmetic.
    // Both lists' roots are one quadratic, a s^2 + b s + c, with the kind's
    // coefficients (a sphere: a = 1 and the eps widening of the
    // discriminant), so a wave whose lanes advance different lists forms one
    // square root and one quotient per lane, not both kinds' in turn.
    __device__ void next_entry(bool th, int k, double &ths, int &thit, double &rs, double &rr,
                               int &rix) const {
        double a = 1.0, b = 0.0, c = -1.0;       // an exhausted list's stand-in, overridden
        double r = 0.0, kept = 0.0;
        bool minus = true, widen = false, use_kept = false;
        int itheta = 0, ix = 0;
        const bool valid = th ? k <= ith_amount : k <= ir_amount;
        if (th) {
            if (valid) {
                int is, root;
                th_index(k, is, root, itheta);
                if (iy0 > 0) {
                    const double tanth2 = tan2[iy0 + is - 1];
                    a = tanth2 * costh02 - sinth02;
                    b = 2.0 * tanth2 * costh0 * z0;
                    c = tanth2 * z0 * z0 - x0 * x0;
                    minus = (root == 1) == (a > 0.0);
                } else {
                    kept = root == 1 ? sar1[(size_t)is * stride] : sar2[(size_t)is * stride];
                    use_kept = true;
                }
            }
        } else if (valid) {
            const int is = k <= half ? k : ir_amount + 1 - k;
            ix = g.nr - is + 1;
            r = ix == 0 ? g.rstar : g.R(ix);
            b = 2.0 * costh0 * z0;
            c = x0 * x0 + z0 * z0 - r * r;
            minus = k <= half;
            widen = true;
        }
        double disc = b * b - 4.0 * a * c;
        if (widen) disc = disc + TELESC_EPS * b * b;
        const double sd = sqrt(disc);
        double s = minus ? (-b - sd) / (2.0 * a) : (-b + sd) / (2.0 * a);
        if (use_kept) s = kept;
        if (!valid) s = 1.0e30;
        if (th) {
            ths = s;
            thit = itheta;
        } else {
            rs = s;
            rr = r;
            rix = ix;
        }
    }
};

// Builds the crossing list and streams it to sink(k, crossing).
// Returns the node count (>= 0) or minus a stop code.
template <typename Sink>
__device__ int trajectory(Traj &t, Sink &&sink) {
    const GridView &g = t.g;
    const double eps = TELESC_EPS, epsplus = 1.0e1 * TELESC_EPS;
    const double pi = PICONST, send = 1.0e30;
    const int nr = g.nr, ny = g.ny;
    (void)ny;
    const double x0 = t.x0, z0 = t.z0;

    t.pitheta0 = 0.5 * pi - t.theta0;
    t.costh02 = t.costh0 * t.costh0;
    t.sinth02 = t.sinth0 * t.sinth0;
    const double costh0 = t.costh0, sinth02 = t.sinth02, costh02 = t.costh02;
    t.iyeq = ny / 2;
    t.iyend = ny;

    // theta = const cones (telescope.F:2968-3000): which have two roots
    int is = 1, first = 0;
    bool contiguous = true;
    for (int iy = 1; iy <= t.iyeq; iy++) {
        double tanth2 = t.tan2[iy];
        double a = tanth2 * costh02 - sinth02;
        double b = 2.0 * tanth2 * costh0 * z0;
        double c = tanth2 * z0 * z0 - x0 * x0;
        double sd = b * b - 4.0 * a * c;
        if (sd > 0.0) {
            if (first == 0) first = iy;
            else if (iy != first + is - 1) contiguous = false;
            is++;
        }
    }
    t.isnr = is - 1;
    if (contiguous && t.isnr > 0) {
        t.iy0 = first;
    } else {
        // kept roots (the same arithmetic)
        is = 1;
        for (int iy = 1; iy <= t.iyeq; iy++) {
            double tanth2 = t.tan2[iy];
            double a = tanth2 * costh02 - sinth02;
            double b = 2.0 * tanth2 * costh0 * z0;
            double c = tanth2 * z0 * z0 - x0 * x0;
            double sd = b * b - 4.0 * a * c;
            if (sd > 0.0) {
                sd = sqrt(sd);
                double s1 = (-b - sd) / (2.0 * a), s2 = (-b + sd) / (2.0 * a);
                if (s1 > s2) { double d = s1; s1 = s2; s2 = d; }
                t.iyar[(size_t)is * t.stride] = iy;
                t.sar1[(size_t)is * t.stride] = s1;
                t.sar2[(size_t)is * t.stride] = s2;
                is++;
            }
        }
    }
    t.isdblnr = 0;
    for (is = 1; is <= t.isnr; is++) {
        double pit = 0.5 * pi - g.TH(t.IY(is));
        if (fabs(pit) > fabs(t.pitheta0)) t.isdblnr++;
    }
    t.iup = t.pitheta0 > 0.0 ? 1 : 0;
    t.branchA = z0 * t.pitheta0 > 0.0 ? 1 : 0;
    t.ith_amount = 2 * t.isnr;

    // R = const spheres (telescope.F:3241-3290)
    double bimpact = sqrt(x0 * x0 + z0 * z0 * (1.0 - costh02));
    int ix;
    for (ix = 1; ix <= nr; ix++)
        if (g.R(ix) > bimpact) break;
    if (ix > nr) return -E_INPUT;
    t.ir_min = bimpact <= g.rstar ? 0 : ix;
    t.bimpact = bimpact;
    t.ir_amount = 2 * (nr + 1 - t.ir_min);
    t.half = nr + 1 - t.ir_min;
    {   // discriminant sanity (telescope.F:3293-3300)
        double r = g.R(nr), b = 2.0 * costh0 * z0, c = x0 * x0 + z0 * z0 - r * r;
        if (b * b - 4.0 * c + eps * b * b < 0.0) return -E_INPUT;
    }

    // extra points (telescope.F:3417-3606).  An extra point is its s and a
    // code (idx + 1) << 1 | rext; its r is formed again where the merge
    // reaches it (the impact parameter for the R-extremum, rext; thr(s) for
    // every other point, as it was formed here).
    // The reference sorts them (ray_sort, nrecip.F:833: an insertion sort,
    // stable).  Each group -- the R-extremum and the 4 x RAYADPT points of
    // its four intervals, then the theta-extremum and its points -- is
    // written here already in that order: with the intervals' ends in
    // order, the points of interval 4 lie below those of 2, then the
    // extremum, then 1, then 3, each interval's points ascending.  Points of
    // equal s and equal code are interchangeable; the R-extremum (rext) goes
    // before the points of interval 2 that equal it, as the stable sort puts
    // it.  A first group whose interval ends are not in that order (the
    // outermost rings, whose lowest interval starts at the r_s(0) = 0 of the
    // reference build) is sorted as the reference sorts it.  The merge below
    // takes the two sorted groups in turn, the first group's point first on
    // equal s, which is the order of the stable sort of both.
    double ex_s[EXMAX + 1];
    int ex_c[EXMAX + 1];
    int nex = 0, n1 = 0;
    auto put = [&](double v, int c) {
        nex++;
        ex_s[nex] = v;
        ex_c[nex] = c;
    };
    if (bimpact > g.R(1)) {
        const double s3 = 0.0 - z0 * costh0;
        if (nex + 1 + 4 * RAYADPT > EXMAX) return -E_EXOVF;
        const int code_ext = (t.ir_min << 1) | 1;          // idx ir_min - 1, rext
        const int isrt = hunt_bisect([&](int i) { return t.r_s(i); }, t.ir_amount, s3);
        // the intervals irng = 1..4 of the reference: (s3, rB), (rA, s3), (rB, rC), (rZ, rA)
        const double rZ = t.r_s(isrt - 1), rA = t.r_s(isrt), rB = t.r_s(isrt + 1),
                     rC = t.r_s(isrt + 2);
        auto points = [&](double s1, double s2) {
            const double ds = (s2 - s1) / (1.0 + 1.0 * RAYADPT);
            for (int iad = 1; iad <= RAYADPT; iad++) put(iad * ds + s1, 1 << 1);
        };
        if (isrt >= 2 && isrt + 2 <= t.ir_amount && rZ <= rA && rA < s3 && s3 <= rB &&
            rB <= rC) {
            points(rZ, rA);                                 // irng 4
            {                                               // irng 2, split at s3
                const double ds = (s3 - rA) / (1.0 + 1.0 * RAYADPT);
                int iad = 1;
                for (; iad <= RAYADPT; iad++) {
                    const double s0 = iad * ds + rA;
                    if (!(s0 < s3)) break;
                    put(s0, 1 << 1);
                }
                put(s3, code_ext);
                for (; iad <= RAYADPT; iad++) put(iad * ds + rA, 1 << 1);
            }
            points(s3, rB);                                 // irng 1
            points(rB, rC);                                 // irng 3
        } else {
            put(s3, code_ext);
            points(s3, rB);
            points(rA, s3);
            points(rB, rC);
            points(rZ, rA);
            // ray_sort (nrecip.F:833): ascending in s, insertion sort
            for (int i = 2; i <= nex; i++) {
                const double ks = ex_s[i];
                const int kc = ex_c[i];
                int j;
                for (j = i - 1; j >= 1 && ex_s[j] > ks; j--) {
                    ex_s[j + 1] = ex_s[j];
                    ex_c[j + 1] = ex_c[j];
                }
                ex_s[j + 1] = ks;
                ex_c[j + 1] = kc;
            }
        }
    }
    n1 = nex;
    {
        const double s3 = x0 * x0 * costh0 / (z0 * sinth02);
        const double rr = t.thr(s3);
        if (rr > g.R(1) && rr < g.R(nr)) {
            int ixh = t.hunt_r(rr);
            if (ixh == 0 || ixh == nr) return -E_HUNT;
            if (nex + 1 + 4 * RAYADPT > EXMAX) return -E_EXOVF;
            const int isrt = hunt_bisect([&](int i) { return t.th_s(i); }, t.ith_amount, s3);
            if (!(isrt - RAYRNPT + 1 < 1 || isrt + RAYRNPT > t.ith_amount)) {
                const double tZ = t.th_s(isrt - 1), tA = t.th_s(isrt), tB = t.th_s(isrt + 1),
                             tC = t.th_s(isrt + 2);
                // the reference's checks, in its interval order irng = 1..4;
                // passed, they put the ends in order: tZ < tA < s3 < tB < tC
                const double e1[4] = {s3, tA, tB, tZ}, e2[4] = {tB, s3, tC, tA};
                for (int q = 0; q < 4; q++) {
                    if (e2[q] == e1[q]) return -E_EXEQ;
                    if (e2[q] < e1[q]) return -E_EXORD;
                }
                auto points = [&](double s1, double s2) {
                    const double ds = (s2 - s1) / (1.0 + 1.0 * RAYADPT);
                    for (int iad = 1; iad <= RAYADPT; iad++) {
                        const double s0 = iad * ds + s1, r2 = t.thr(s0);
                        const int ixx = t.hunt_r(r2);
                        if (ixx == 0 || ixx == nr) continue;        // :3580
                        put(s0, 1 << 1);                            // r = r2 = thr(s0)
                    }
                };
                points(tZ, tA);                                     // irng 4
                points(tA, s3);                                     // irng 2
                put(s3, 1 << 1);                                    // r = rr = thr(s3)
                points(s3, tB);                                     // irng 1
                points(tB, tC);                                     // irng 3
            } else {
                put(s3, 1 << 1);
            }
        }
    }

    // merge (telescope.F:3610-3677)
    const double rmaxt = g.R(nr) * (1.0 - epsplus), rmaxr = g.R(nr) * (1.0 + epsplus);
    const double rmint = g.R(1) * (1.0 + epsplus), rminr = g.R(1) * (1.0 - epsplus);
    const double sbeg = -1.0e30;
    double sprev = -1.0e30;
    int nout = 0, ist = 1, isr = 1;
    double ths, rs, rr_;
    int thit, rix;
    t.th_entry(ist, ths, thit);
    t.r_entry(isr, rs, rr_, rix);
    auto emit = [&](int icr, double s, int idx, int rext) {
        Crossing c;
        c.s = s; c.icross = icr; c.idx = idx; c.rext = rext;
        sink(nout, c);
        nout++;
        sprev = s;
    };
    const int tr_total = t.ir_amount + t.ith_amount;
    // the next point of each group of extra points, kept in registers
    int i1 = 1, i2 = n1 + 1;
    double es1 = n1 >= 1 ? ex_s[1] : 1.0e30, es2 = nex > n1 ? ex_s[n1 + 1] : 1.0e30;
    for (int iss = 1; iss <= tr_total; iss++) {
        for (;;) {
            const bool second = es2 < es1;
            const double es = second ? es2 : es1;
            double m = ths < rs ? ths : rs;
            if (!(es < m)) break;
            const int code = ex_c[second ? i2 : i1];
            const double er = (code & 1) ? t.bimpact : t.thr(es);
            if (er <= rmaxt && er >= rmint && (es - sprev) > eps * er && es >= sbeg &&
                es <= send)
                emit(3, es, (code >> 1) - 1, code & 1);
            if (second) {
                i2++;
                es2 = i2 <= nex ? ex_s[i2] : 1.0e30;
            } else {
                i1++;
                es1 = i1 <= n1 ? ex_s[i1] : 1.0e30;
            }
        }
        // the nearer of the next cone and the next sphere crossing
        const bool th = ths < rs;
        double r = rr_;
        if (th) r = t.thr(ths);
        const double s = th ? ths : rs;
        const bool inside = th ? (r <= rmaxt && r >= rmint) : (r <= rmaxr && r >= rminr);
        if (inside && (s - sprev) > eps * r && s >= sbeg && s <= send)
            emit(th ? 2 : 1, s, th ? thit : rix, 0);
        if (th) ist++;
        else isr++;
        t.next_entry(th, th ? ist : isr, ths, thit, rs, rr_, rix);
    }
    return nout;
}

}  // namespace rl
