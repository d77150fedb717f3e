// sc_pair.h -- device-side pair energy of patchy spherocylinders and spheres.
//
// The physics is that of robertvacha/SC's PairE functors (scOOP/mc/paire.h,
// paire.cpp); every branch whose outcome defines an energy value (the 0.0
// "empty" sentinel of intersection lists, exact-equality de-duplication,
// clamping order in the segment distance) is kept as the reference has it, so
// that FP64 instantiations agree with the reference to rounding.  The code is
// organised for the GPU:
//   * one templated implementation, R = float (the product path) or double
//     (the FP64 cross-check mode and the near-parallel fallback);
//   * per-particle derived vectors (patch sides, chiral axes) are computed once
//     when a particle is staged, not per pair;
//   * a pair is split into a cheap stage (minimum image, cutoff, segment
//     distance, WCA) and an expensive stage (patch-wedge intersections and the
//     attraction), so kernels can compact the expensive work across lanes.
#pragma once

// FMA contraction only within one expression (C's FP_CONTRACT ON), never
// across statements: with cross-statement fusion the rounding of a derived
// vector or a pair term depends on what the optimiser sees around the inlined
// call, so the sweep's trial energies and the whole-system energy could round
// the same pair differently and, where an FP32 branch sits on a knife edge
// (an intersection test or de-duplication), take different branches.  With
// per-expression contraction every kernel computes a pair bit for bit alike.
// The pragma stays in force for the rest of the including file.
#pragma clang fp contract(on)

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SC_HD __host__ __device__ __forceinline__

namespace scg {

// geometric type codes, scOOP/structures/macros.h:71-85
enum : int {
    G_SCN = 10, G_SCA = 11, G_PSC = 12, G_CPSC = 13, G_CHPSC = 14, G_CHCPSC = 15, G_TPSC = 16,
    G_TCPSC = 17, G_TCHPSC = 18, G_TCHCPSC = 19, G_SPN = 30, G_SPA = 31
};
SC_HD bool is_psc(int g) { return g == G_PSC || g == G_CHPSC || g == G_TPSC || g == G_TCHPSC; }
SC_HD bool is_cpsc(int g) { return g == G_CPSC || g == G_CHCPSC || g == G_TCPSC || g == G_TCHCPSC; }
SC_HD bool is_chiral(int g) { return g == G_CHPSC || g == G_CHCPSC || g == G_TCHPSC || g == G_TCHCPSC; }
SC_HD bool is_two(int g) { return g == G_TPSC || g == G_TCPSC || g == G_TCHPSC || g == G_TCHCPSC; }

// functor kinds selected per type pair (PairE::initIntFCE, paire.cpp:6-80)
enum : int {
    K_NONE = -1, K_PSC = 0, K_CPSC = 1, K_MIXED = 2, K_SCN = 3, K_SCA = 4,
    K_SPH_WCA = 10, K_SPH_TAYLOR = 11, K_SCASPA = 20, K_PSCSPA = 21, K_CPSCSPA = 22
};

// Pair parameters for one ordered (dense) type pair.  FP64 at the host, both
// precisions on the device.
template <typename R>
struct PairParamsT {
    int kind, g0, g1, exclude;
    int fT, sT;              // first / second particle carries a second patch
    int isp1;                // MixSpSc: particle 1 is the spherocylinder
    int sp_chiral, sp_sec;   // MixSpSc: getGeoTypes of the (transposed) parameter
    R hl0, hl1;
    R eps, A, B, pdis, pswitch, pswitchINV, rcut, rcutSq, rcutwca, rcutwcaSq;
    R pcangl[4], pcanglsw[4];
    R parallel;
    // parameters of the transposed pair used by MixSpSc when particle 2 is the SC
    R t_hl0, t_eps, t_pdis, t_pswitch, t_rcut, t_rcutSq, t_rcutwcaSq, t_pcangl0, t_pcanglsw0;
    int t_exclude, pad0;
};

// Per-type constants for the derived vectors (Particle::init, particle.cpp:3-79).
template <typename R>
struct TypeParamsT {
    int geotype, pad;
    R pcoshalfi0, psinhalfi0, pcoshalfi2, psinhalfi2;
    R csec, ssec, chcos, chsin;
};

template <typename R>
struct V3 {
    R x, y, z;
};
template <typename R> SC_HD V3<R> mk(R x, R y, R z) { V3<R> v; v.x = x; v.y = y; v.z = z; return v; }
template <typename R> SC_HD R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> SC_HD V3<R> scl(R s, const V3<R>& a) { return mk<R>(a.x * s, a.y * s, a.z * s); }
template <typename R> SC_HD V3<R> sub(const V3<R>& a, const V3<R>& b) { return mk<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> SC_HD V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return mk<R>(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
template <typename R> SC_HD V3<R> perp(const V3<R>& a, const V3<R>& b) {
    R dp = dot(a, b);
    return mk<R>(a.x - b.x * dp, a.y - b.y * dp, a.z - b.z * dp);
}
template <typename R> SC_HD R len(const V3<R>& a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
template <typename R> SC_HD V3<R> normalised(V3<R> a) {
    R t = len(a);
    if (t != R(0)) { t = R(1) / t; a.x *= t; a.y *= t; a.z *= t; }
    return a;
}
// quaternion rotation by (c, s*ax) -- Vector::rotate, Vector.h:138-160
template <typename R> SC_HD V3<R> qrot(const V3<R>& v, const V3<R>& ax, R c, R s) {
    R qx = ax.x * s, qy = ax.y * s, qz = ax.z * s;
    R t2 = c * qx, t3 = c * qy, t4 = c * qz, t5 = -qx * qx, t6 = qx * qy, t7 = qx * qz, t8 = -qy * qy, t9 = qy * qz,
      t10 = -qz * qz;
    return mk<R>(R(2) * ((t8 + t10) * v.x + (t6 - t4) * v.y + (t3 + t7) * v.z) + v.x,
                 R(2) * ((t4 + t6) * v.x + (t5 + t10) * v.y + (t9 - t2) * v.z) + v.y,
                 R(2) * ((t7 - t3) * v.x + (t2 + t9) * v.y + (t5 + t8) * v.z) + v.z);
}

// A particle with its derived vectors.  ax[k] is the axis patch k is attached
// to: chdir[k] for chiral types, dir otherwise (the CH flags of every SC-SC and
// SC-sphere functor are exactly "the particle's geotype is chiral").
template <typename R>
struct DPart {
    V3<R> dir;
    V3<R> pd[2];
    V3<R> sd[4];
    V3<R> ax[2];
    int t;      // dense type
    int g;      // geotype
};

template <typename R>
SC_HD void derive(DPart<R>& p, const TypeParamsT<R>& tp) {
    int g = tp.geotype;
    p.g = g;
    p.ax[0] = p.dir;
    p.ax[1] = p.dir;
    if (g >= 30 || g == G_SCA || g == G_SCN) return;
    p.dir = normalised(p.dir);
    p.pd[0] = normalised(perp(p.pd[0], p.dir));
    p.ax[0] = p.dir;
    p.ax[1] = p.dir;
    if (is_two(g)) {
        p.pd[1] = normalised(perp(qrot(p.pd[0], p.dir, tp.csec, tp.ssec), p.dir));
    }
    if (is_chiral(g)) {
        p.ax[0] = qrot(p.dir, p.pd[0], tp.chcos, tp.chsin);
        if (is_two(g)) p.ax[1] = qrot(p.dir, p.pd[1], tp.chcos, tp.chsin);
    }
    p.sd[0] = qrot(p.pd[0], p.ax[0], tp.pcoshalfi0, tp.psinhalfi0);
    p.sd[1] = qrot(p.pd[0], p.ax[0], tp.pcoshalfi0, -tp.psinhalfi0);
    if (is_two(g)) {
        p.sd[2] = qrot(p.pd[1], p.ax[1], tp.pcoshalfi2, tp.psinhalfi2);
        p.sd[3] = qrot(p.pd[1], p.ax[1], tp.pcoshalfi2, -tp.psinhalfi2);
    }
}

// ------------------------------------------------------------------ segments
// Closest vector between two centred segments, EPatch::minDistSegments
// (paire.cpp:86-241): clamped (s,t) solve, then for near-parallel lines
// (D < 1e-8) the same solve with the roles swapped, keeping the shorter.
template <typename R>
SC_HD V3<R> clamp_solve(const V3<R>& u, const V3<R>& v, const V3<R>& w, R aa, R b, R cc, R d, R e, R D, bool first,
                        bool& par) {
    R sN = D, sD = D, tN = D, tD = D;
    if (D < R(1e-8)) {
        par = true;
        sN = R(0); sD = R(1); tN = e; tD = cc;
    } else if (first) {
        sN = b * e - cc * d;
        tN = aa * e - b * d;
        if (sN < R(0)) { sN = R(0); tN = e; tD = cc; }
        else if (sN > sD) { sN = sD; tN = e + b; tD = cc; }
    }
    if (tN < R(0)) {
        tN = R(0);
        if (-d < R(0)) sN = R(0);
        else if (-d > aa) sN = sD;
        else { sN = -d; sD = aa; }
    } else if (tN > tD) {
        tN = tD;
        R q = -d + b;
        if (q < R(0)) sN = R(0);
        else if (q > aa) sN = sD;
        else { sN = q; sD = aa; }
    }
    R sc = (fabs(sN) < R(1e-8)) ? R(0) : sN / sD;
    R tc = (fabs(tN) < R(1e-8)) ? R(0) : tN / tD;
    return mk<R>(u.x * sc + w.x - v.x * tc, u.y * sc + w.y - v.y * tc, u.z * sc + w.z - v.z * tc);
}

template <typename R>
SC_HD V3<R> min_dist_impl(const V3<R>& A, const V3<R>& Bv, R h1, R h2, const V3<R>& r) {
    V3<R> u = scl(R(2) * h1, A), v = scl(R(2) * h2, Bv);
    V3<R> w = mk<R>(Bv.x * h2 - A.x * h1 - r.x, Bv.y * h2 - A.y * h1 - r.y, Bv.z * h2 - A.z * h1 - r.z);
    R a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    R D = a * c - b * b;
    bool par = false;
    V3<R> vec = clamp_solve(u, v, w, a, b, c, d, e, D, true, par);
    if (par) {
        V3<R> w2 = mk<R>(A.x * h1 - Bv.x * h2 + r.x, A.y * h1 - Bv.y * h2 + r.y, A.z * h1 - Bv.z * h2 + r.z);
        bool dummy = false;
        V3<R> vec2 = clamp_solve(v, u, w2, c, b, a, dot(v, w2), dot(u, w2), D, false, dummy);
        if (dot(vec2, vec2) < dot(vec, vec)) return vec2;
    }
    return vec;
}

// FP32 entry: when the lines are close to parallel the FP32 determinant
// a*c-b*b loses all its digits (for L=3, a=c=9 and D=81 sin^2), so the whole
// solve, including the reference's D<1e-8 decision, is done in FP64.  The
// switch is taken on |u x v|^2 (accurate in FP32) relative to a*c.
SC_HD V3<float> min_dist(const V3<float>& A, const V3<float>& Bv, float h1, float h2, const V3<float>& r) {
    V3<float> cr = cross(A, Bv);
    float s2 = dot(cr, cr);   // sin^2 of the angle between the unit axes
    float hh = 4.0f * h1 * h2;
    if (s2 < 2e-3f || hh * hh * s2 < 1e-6f) {   // relative or absolute D small
        V3<double> Ad = mk<double>(A.x, A.y, A.z), Bd = mk<double>(Bv.x, Bv.y, Bv.z), rd = mk<double>(r.x, r.y, r.z);
        V3<double> o = min_dist_impl<double>(Ad, Bd, (double)h1, (double)h2, rd);
        return mk<float>((float)o.x, (float)o.y, (float)o.z);
    }
    return min_dist_impl<float>(A, Bv, h1, h2, r);
}
SC_HD V3<double> min_dist(const V3<double>& A, const V3<double>& Bv, double h1, double h2, const V3<double>& r) {
    return min_dist_impl<double>(A, Bv, h1, h2, r);
}

// ---------------------------------------------------------------- patches
template <typename R> SC_HD R fangl(R a, R pc, R pcsw) {     // fanglScale, paire.h:12-17
    if (a <= pcsw) return R(0);
    return (a >= pc) ? R(1) : (R(0.5) - ((pcsw + pc) * R(0.5) - a) / (pc - pcsw));
}

// Intersection parameter list with the reference's "0.0 means empty" rule
// (testIntrPatch, paire.h:86-101).
template <typename R>
struct Ints {
    R i0, i1;
    SC_HD void add_patch(const V3<R>& dir, const V3<R>& pdir, V3<R> vec, R cospatch, R ti) {
        vec = perp(vec, dir);
        if (dot(pdir, vec) >= cospatch * len(vec)) {
            if (i0 == R(0)) { i0 = ti; return; }
            if (i1 == R(0) && i0 != ti) i1 = ti;
        }
    }
};

// Everything one directional intersection query needs: segment 2 (axis p2,
// half length h2, centre at -r relative to particle 1) against the attractive
// region of patch (pdir, s0, s1) on particle 1 (axis p1, half length h1).
template <typename R>
struct IQ {
    V3<R> p1, p2, pdir, s0, s1, r;
    R cossw, rcSq, h1, h2;
};

// findIntersectPlaneUni, paire.h:150-181
template <typename R>
SC_HD bool plane_hit(const IQ<R>& q, const V3<R>& side, R& ti, R& c, R& d) {
    V3<R> np = cross(q.p1, side);
    R a = dot(np, q.p2);
    c = R(1); d = R(1);
    if (a == R(0)) return false;
    ti = dot(np, q.r) / a;
    if ((ti > q.h2) || (ti < -q.h2)) return false;
    V3<R> dv = mk<R>(ti * q.p2.x - q.r.x, ti * q.p2.y - q.r.y, ti * q.p2.z - q.r.z);
    c = dot(dv, side);
    if (c * q.cossw < R(0)) return false;
    d = fabs(dot(dv, q.p1)) - q.h1;
    return true;
}

// point x*p2 - r of segment 2; `inside` selects scToInfiIntr (inside the
// cylinder slab, paire.h:103-119) or scToEndSpIntr (outside it, :588-602)
template <typename R>
SC_HD void line_point(const IQ<R>& q, Ints<R>& in, R x, bool inside) {
    if ((x >= q.h2) || (x <= -q.h2)) return;
    V3<R> v = mk<R>(q.p2.x * x - q.r.x, q.p2.y * x - q.r.y, q.p2.z * x - q.r.z);
    R e = dot(q.p1, v);
    bool out = (e >= q.h1) || (e <= -q.h1);
    if (out != inside) in.add_patch(q.p1, q.pdir, v, q.cossw, x);
}

// testIntrAtC, paire.h:121-148: infinite cylinder of radius rcut
template <typename R>
SC_HD void cyl_hits(const IQ<R>& q, Ints<R>& in) {
    V3<R> v1 = cross(mk<R>(R(-1) * q.r.x, R(-1) * q.r.y, R(-1) * q.r.z), q.p1), v2 = cross(q.p2, q.p1);
    R a = dot(v2, v2), b = R(2) * dot(v1, v2), c = -q.rcSq + dot(v1, v1), d = b * b - R(4) * a * c;
    if (d >= R(0)) {
        d = sqrt(d);
        a = R(0.5) / a;
        line_point(q, in, (-b + d) * a, true);
        if (d > R(0)) line_point(q, in, (-b - d) * a, true);
    }
}

// Psc::pscIntersect (paire.h:499-583) and CPsc::cpscIntersect (:687-826):
// returns true with (t1,t2) when segment 2 enters and leaves the region.
template <typename R>
SC_HD bool intersect(const IQ<R>& q, bool cyl_patch, R& t1, R& t2) {
    Ints<R> in;
    in.i0 = R(0); in.i1 = R(0);
    R ti, c, d;
    // 1a: the two half-planes bounding the patch wedge
    for (int k = 0; k < 2; k++) {
        if (plane_hit(q, k == 0 ? q.s0 : q.s1, ti, c, d)) {
            bool ok = cyl_patch ? (d <= R(0) && c * c <= q.rcSq) : (((d <= R(0)) ? c * c : d * d + c * c) <= q.rcSq);
            if (ok) {
                if (k == 0) in.i0 = ti;
                else if (in.i0 == R(0)) in.i0 = ti;
                else if (ti != in.i0) in.i1 = ti;
            }
        }
    }
    if (in.i1 != R(0)) { t1 = in.i0; t2 = in.i1; return true; }
    // 1b: the cylindrical surface at distance rcut
    cyl_hits(q, in);
    if (in.i1 != R(0)) { t1 = in.i0; t2 = in.i1; return true; }
    // 1c: the end caps -- spheres for PSC, flat plates for CPSC
    if (!cyl_patch) {
        for (int k = 0; k < 2; k++) {
            R sg = (k == 0) ? R(1) : R(-1);
            V3<R> vc = mk<R>(sg * q.p1.x * q.h1 - q.r.x, sg * q.p1.y * q.h1 - q.r.y, sg * q.p1.z * q.h1 - q.r.z);
            R b = R(2) * dot(vc, q.p2), cq = dot(vc, vc) - q.rcSq, dd = b * b - R(4) * cq;
            if (dd >= R(0)) {
                dd = sqrt(dd);
                line_point(q, in, (-b + dd) * R(0.5), false);
                if (dd > R(0)) line_point(q, in, (-b - dd) * R(0.5), false);
            }
        }
    } else {
        R a = dot(q.p1, q.p2);
        if (a != R(0)) {
            for (int k = 0; k < 2; k++) {
                V3<R> v1 = (k == 0) ? mk<R>(q.r.x + q.h1 * q.p1.x, q.r.y + q.h1 * q.p1.y, q.r.z + q.h1 * q.p1.z)
                                    : mk<R>(q.r.x - q.h1 * q.p1.x, q.r.y - q.h1 * q.p1.y, q.r.z - q.h1 * q.p1.z);
                R x = dot(q.p1, v1) / a;
                if (!((x > q.h2) || (x < -q.h2))) {
                    V3<R> v2 = mk<R>(x * q.p2.x - v1.x, x * q.p2.y - v1.y, x * q.p2.z - v1.z);
                    if (!(dot(v2, v2) > q.rcSq)) in.add_patch(q.p1, q.pdir, v2, q.cossw, x);
                }
            }
        }
    }
    if (in.i1 != R(0)) { t1 = in.i0; t2 = in.i1; return true; }
    // 1d: an end of segment 2 inside the region (testIntrA, paire.h:625-652;
    // cpscIntersect :786-821 accepts only the cylindrical part)
    for (int k = 0; k < 2 && in.i1 == R(0); k++) {
        R hh = (k == 0) ? q.h2 : -q.h2;
        V3<R> v1 = mk<R>(q.p2.x * hh - q.r.x, q.p2.y * hh - q.r.y, q.p2.z * hh - q.r.z);
        R a = dot(v1, q.p1);
        V3<R> v2 = mk<R>(v1.x - q.p1.x * a, v1.y - q.p1.y * a, v1.z - q.p1.z * a);
        R b = dot(v2, v2), dd = fabs(a) - q.h1;
        bool ok = cyl_patch ? (dd <= R(0) && b < q.rcSq) : (((dd <= R(0)) ? b : dd * dd + b) < q.rcSq);
        if (ok) in.add_patch(q.p1, q.pdir, v1, q.cossw, cyl_patch ? (k == 0 ? q.h2 : R(-1) * q.h2) : hh);
    }
    t1 = in.i0; t2 = in.i1;
    return in.i1 != R(0);
}

template <typename R> SC_HD R cos_well(R dist, R pdis, R pswitch, R eps) {
    if (dist < pdis) return -eps;
    R e = cos(R(1.57079632679489661923132169163975) * (dist - pdis) / pswitch);
    return e * (-e * eps);
}

// One patch of one particle as seen by the patch-pair functor: the axis the
// patch is attached to, the patch direction and the two side vectors.
template <typename R>
struct PatchView {
    V3<R> ax, pd, s0, s1;
};

template <typename R>
SC_HD PatchView<R> view(const DPart<R>& p, int k) {
    PatchView<R> v;
    v.ax = k ? p.ax[1] : p.ax[0];
    v.pd = k ? p.pd[1] : p.pd[0];
    v.s0 = k ? p.sd[2] : p.sd[0];
    v.s1 = k ? p.sd[3] : p.sd[1];
    return v;
}

// One patch pair of two spherocylinders: Psc/CPsc/PscCPsc::operator()
// (paire.h:469-484, :658-672, :864-889) and EPatch::atrE (paire.cpp:244-301).
// k1/k2 select the patch of each particle; they only pick parameter indices,
// so one code path serves all four patch pairs (no divergence between lanes
// evaluating different patch pairs).  The pair term is two independent
// wedge intersections (patch_side 0: particle 2's axis against particle 1's
// patch region, giving T1, T2; side 1: the reverse, S1, S2), zero unless both
// hit, then the attraction of the overlap (patch_atr).
template <typename R>
SC_HD bool patch_side(const PairParamsT<R>& P, const PatchView<R>& A, const PatchView<R>& Bv, const V3<R>& r, int k1,
                      int k2, int side, R& x1, R& x2) {
    bool c1 = (P.kind == K_CPSC) || (P.kind == K_MIXED && !is_psc(P.g0));
    bool c2 = (P.kind == K_CPSC) || (P.kind == K_MIXED && is_psc(P.g0));
    R sw1 = k1 ? P.pcanglsw[2] : P.pcanglsw[0];
    R sw2 = k2 ? P.pcanglsw[3] : P.pcanglsw[1];
    // one intersect call for either side (lanes of both sides share its path)
    bool s1 = side != 0;
    IQ<R> q;
    q.rcSq = P.rcutSq;
    q.p1 = s1 ? Bv.ax : A.ax;
    q.p2 = s1 ? A.ax : Bv.ax;
    q.pdir = s1 ? Bv.pd : A.pd;
    q.s0 = s1 ? Bv.s0 : A.s0;
    q.s1 = s1 ? Bv.s1 : A.s1;
    q.r = s1 ? mk<R>(r.x * R(-1), r.y * R(-1), r.z * R(-1)) : r;
    q.cossw = s1 ? sw2 : sw1;
    q.h1 = s1 ? P.hl1 : P.hl0;
    q.h2 = s1 ? P.hl0 : P.hl1;
    return intersect(q, s1 ? c2 : c1, x1, x2);
}

template <typename R>
SC_HD R patch_atr(const PairParamsT<R>& P, const PatchView<R>& A, const PatchView<R>& Bv, const V3<R>& r, int k1,
                  int k2, R T1, R T2, R S1, R S2) {
    R sw1 = k1 ? P.pcanglsw[2] : P.pcanglsw[0], pc1 = k1 ? P.pcangl[2] : P.pcangl[0];
    R sw2 = k2 ? P.pcanglsw[3] : P.pcanglsw[1], pc2 = k2 ? P.pcangl[3] : P.pcangl[1];
    const V3<R>& d1 = A.ax;
    const V3<R>& d2 = Bv.ax;
    R v1 = fabs(S1 - S2), v2 = fabs(T1 - T2), f0 = R(0.5) * (v1 + v2);
    R m1 = (S1 + S2) * R(0.5), m2 = (T1 + T2) * R(0.5);
    V3<R> vi = mk<R>(d2.x * m2 - d1.x * m1 - r.x, d2.y * m2 - d1.y * m1 - r.y, d2.z * m2 - d1.z * m1 - r.z);
    V3<R> vm = min_dist(d1, d2, v1, v2, vi);
    R e = cos_well<R>((R)sqrt(dot(vm, vm)), P.pdis, P.pswitch, P.eps);
    V3<R> w = perp(vi, d1);
    R f1 = fangl(dot(w, A.pd) / len(w), pc1, sw1);
    w = perp(mk<R>(vi.x * R(-1), vi.y * R(-1), vi.z * R(-1)), d2);
    R f2 = fangl(dot(w, Bv.pd) / len(w), pc2, sw2);
    R paral = R(1);
    if (P.parallel != R(0)) {
        R ca = dot(d1, d2);
        if ((P.parallel > R(0) && ca > R(0)) || (P.parallel < R(0) && ca < R(0))) paral = R(1) + P.parallel * ca;
    }
    return e * (f0 * f1 * f2 * paral);
}

template <typename R>
SC_HD R patch_pair_v(const PairParamsT<R>& P, const PatchView<R>& A, const PatchView<R>& Bv, const V3<R>& r, int k1,
                     int k2) {
    R T1, T2, S1, S2;
    if (!patch_side(P, A, Bv, r, k1, k2, 0, T1, T2)) return R(0);
    if (!patch_side(P, A, Bv, r, k1, k2, 1, S1, S2)) return R(0);
    return patch_atr(P, A, Bv, r, k1, k2, T1, T2, S1, S2);
}

template <typename R, int k1, int k2>
SC_HD R patch_pair(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r) {
    return patch_pair_v<R>(P, view(a, k1), view(b, k2), r, k1, k2);
}

// number of patch pairs of an SC-SC pair and the (k1,k2) of the m-th, in the
// order SpheroCylinder::operator() sums them (paire.h:1141-1170)
template <typename R>
SC_HD int npatch_pairs(const PairParamsT<R>& P) {
    return 1 + (P.fT ? 1 : 0) + (P.sT ? 1 : 0) + ((P.fT && P.sT) ? 1 : 0);
}
template <typename R>
SC_HD void patch_pair_index(const PairParamsT<R>& P, int m, int& k1, int& k2) {
    // m = 0 -> (0,0); then (1,0) if fT; then (0,1) if sT; then (1,1)
    k1 = 0; k2 = 0;
    if (m == 0) return;
    if (P.fT) { if (m == 1) { k1 = 1; return; } m--; }
    if (P.sT) { if (m == 1) { k2 = 1; return; } m--; }
    k1 = 1; k2 = 1;
}

// WCA on a squared distance, WcaTruncSq (paire.h:385-393)
template <typename R> SC_HD R wca_sq(R d2, R rcSq, R eps, R A, R B) {
    if (d2 > rcSq) return R(0);
    R i2 = R(1) / d2, i6 = i2 * i2 * i2;
    return eps + A * (i6 * i6) - B * i6;
}

// Cheap stage of an SC-SC pair: segment distance and repulsion.  Returns
// whether the attractive (expensive) stage applies.
template <typename R>
SC_HD bool scsc_cheap(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r, R& e_rep,
                      R& dist2) {
    V3<R> dv = min_dist(a.dir, b.dir, P.hl0, P.hl1, r);
    dist2 = dot(dv, dv);
    e_rep = wca_sq(dist2, P.rcutwcaSq, P.eps, P.A, P.B);
    if (P.kind == K_SCN) return false;
    if (P.kind == K_SCA) {
        // Sca::operator() (paire.h:845-852) adds the WCA of the same distance
        if (!(dist2 > P.rcutSq || P.eps == R(0) || P.exclude)) {
            R dist = sqrt(dist2);
            if (!(dist > P.rcutwca)) {
                R i2 = R(1) / dist2, i6 = i2 * i2 * i2;
                e_rep += P.A * (i6 * i6) - P.B * i6 + P.eps;
            }
        }
        return false;
    }
    return !(dist2 > P.rcutSq || P.eps == R(0) || P.exclude);
}

// Expensive stage of an SC-SC pair: the sum over patch pairs
// (SpheroCylinder::operator(), paire.h:1141-1170).
template <typename R>
SC_HD R scsc_attr(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r) {
    R e = patch_pair<R, 0, 0>(P, a, b, r);
    if (P.fT) e += patch_pair<R, 1, 0>(P, a, b, r);
    if (P.sT) e += patch_pair<R, 0, 1>(P, a, b, r);
    if (P.fT && P.sT) e += patch_pair<R, 1, 1>(P, a, b, r);
    return e;
}

// The same sum in FP64 over the FP32 patch-pair terms.  The FP32 kernels add
// a pair's terms (the cheap stage's repulsion, each patch pair) in FP64: the
// sweep evaluates them on different lanes and can only add them in FP64, so
// every other kernel adds the same FP32 values the same way, and a pair has
// one value to ~1e-16 wherever it is priced.
template <typename R>
SC_HD double scsc_attr_d(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r) {
    double e = (double)patch_pair<R, 0, 0>(P, a, b, r);
    if (P.fT) e += (double)patch_pair<R, 1, 0>(P, a, b, r);
    if (P.sT) e += (double)patch_pair<R, 0, 1>(P, a, b, r);
    if (P.fT && P.sT) e += (double)patch_pair<R, 1, 1>(P, a, b, r);
    return e;
}

// MixSpSc::closestDist, paire.h:1077-1093
template <typename R>
SC_HD R sp_dist(const V3<R>& r, const V3<R>& d1, R hl, R& contt, V3<R>& dv) {
    contt = dot(d1, r);
    R d = -hl;
    if (contt >= hl) d = hl;
    else if (contt > -hl) d = contt;
    dv = mk<R>(-r.x + d1.x * d, -r.y + d1.y * d, -r.z + d1.z * d);
    return dot(dv, dv);
}

// PscSpa / ScaSpa::operator(), paire.h:914-979
template <typename R>
SC_HD R patch_sphere(bool sca, R dist, R contt, const V3<R>& dv, R hl, R rcSq, R pdis, R pswitch, R eps, R pc,
                     R pcsw, const V3<R>& axis, const V3<R>& pdir) {
    R e = cos_well<R>(dist, pdis, pswitch, eps);
    R b = (R)sqrt(rcSq - dist * dist);
    R f0 = (contt + b > hl) ? hl : contt + b;
    if (contt - b < -hl) f0 -= -hl;
    else f0 -= contt - b;
    if (sca) return e * f0;
    V3<R> w = perp(dv, axis);
    return e * (fangl(dot(w, pdir) / len(w), pc, pcsw) * f0);
}

// MixSpSc::operator() without bonds, paire.h:1022-1067
template <typename R>
SC_HD R spsc(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r) {
    const DPart<R>& s = P.isp1 ? a : b;
    V3<R> rr = P.isp1 ? r : mk<R>(R(-1) * r.x, R(-1) * r.y, R(-1) * r.z);
    // parameters of iaParam = ia[sc][sphere]
    R hl = P.isp1 ? P.hl0 : P.t_hl0;
    R eps = P.isp1 ? P.eps : P.t_eps, pdis = P.isp1 ? P.pdis : P.t_pdis, psw = P.isp1 ? P.pswitch : P.t_pswitch;
    R rcut = P.isp1 ? P.rcut : P.t_rcut, rcSq = P.isp1 ? P.rcutSq : P.t_rcutSq;
    R rwSq = P.isp1 ? P.rcutwcaSq : P.t_rcutwcaSq;
    R pc = P.isp1 ? P.pcangl[0] : P.t_pcangl0, pcsw = P.isp1 ? P.pcanglsw[0] : P.t_pcanglsw0;
    int excl = P.isp1 ? P.exclude : P.t_exclude;
    R contt = R(0), rep = R(0), atr = R(0);
    V3<R> dv;
    R d2 = sp_dist(rr, s.dir, hl, contt, dv);
    if (d2 < rwSq) rep = wca_sq(d2, P.rcutwcaSq, P.eps, P.A, P.B);
    bool far = (P.kind == K_CPSCSPA) && ((contt > hl) || (contt < -hl));
    if (d2 > rcSq || eps == R(0) || excl || far) return rep;
    bool sca = P.kind == K_SCASPA;
    if (P.sp_chiral) d2 = sp_dist(rr, s.ax[0], hl, contt, dv);
    R dist = (R)sqrt(d2);
    if (dist < rcut) atr = patch_sphere(sca, dist, contt, dv, hl, rcSq, pdis, psw, eps, pc, pcsw, s.ax[0], s.pd[0]);
    if (P.sp_sec) {
        d2 = sp_dist(rr, s.ax[1], hl, contt, dv);
        dist = (R)sqrt(d2);
        if (dist < rcut) atr += patch_sphere(sca, dist, contt, dv, hl, rcSq, pdis, psw, eps, pc, pcsw, s.ax[1], s.pd[1]);
    }
    return rep + atr;
}

// Sphere functors, paire.h:376-383 (WcaTrunc) and :416-439 (WcaCos2Taylor)
template <typename R>
SC_HD R sphere_pair(const PairParamsT<R>& P, R dist) {
    if (P.kind == K_SPH_WCA) {
        if (dist > P.rcutwca) return R(0);
        R i2 = R(1) / (dist * dist), i6 = i2 * i2 * i2;
        return P.A * (i6 * i6) - P.B * i6 + P.eps;
    }
    if (dist > P.rcut || P.eps == R(0) || P.exclude) return R(0);
    R e;
    if (dist > P.pdis) {
        R x = R(1.57079632679489661923132169163975) * (dist - P.pdis) * P.pswitchINV;
        x *= x;
        e = (R(1) - x + x * x * R(1.0 / 3.0) - x * x * x * R(2.0 / 45.0) + x * x * x * x * R(1.0 / 315.0) -
             x * x * x * x * x * R(2.0 / 14175.0) + x * x * x * x * x * x * R(2.0 / 467775.0) -
             x * x * x * x * x * x * x * R(4.0 / 42567525) + x * x * x * x * x * x * x * x * R(1.0 / 638512875)) *
            -P.eps;
    } else {
        e = -P.eps;
    }
    if (dist > P.rcutwca) return e;
    R i2 = R(1) / (dist * dist), i6 = i2 * i2 * i2;
    return P.A * (i6 * i6) - P.B * i6;
}

// Full pair energy (PairE::operator(), paire.h:1209-1220) on a minimum-image
// vector r = pos1 - pos2 already computed by the caller.
template <typename R>
SC_HD R pair_energy(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r, R sqmaxcut) {
    R r2 = dot(r, r);
    if (r2 > sqmaxcut) return R(0);
    switch (P.kind) {
    case K_PSC: case K_CPSC: case K_MIXED: case K_SCN: case K_SCA: {
        R erep, d2;
        bool att = scsc_cheap(P, a, b, r, erep, d2);
        return att ? erep + scsc_attr(P, a, b, r) : erep;
    }
    case K_SPH_WCA: case K_SPH_TAYLOR:
        return sphere_pair(P, (R)sqrt(r2));
    case K_SCASPA: case K_PSCSPA: case K_CPSCSPA:
        return spsc(P, a, b, r);
    default:
        return R(0);
    }
}

// pair_energy with the FP32 terms added in FP64 (scsc_attr_d): the form every
// FP32 kernel uses.
template <typename R>
SC_HD double pair_energy_d(const PairParamsT<R>& P, const DPart<R>& a, const DPart<R>& b, const V3<R>& r,
                           R sqmaxcut) {
    R r2 = dot(r, r);
    if (r2 > sqmaxcut) return 0.0;
    switch (P.kind) {
    case K_PSC: case K_CPSC: case K_MIXED: case K_SCN: case K_SCA: {
        R erep, d2;
        bool att = scsc_cheap(P, a, b, r, erep, d2);
        return att ? (double)erep + scsc_attr_d(P, a, b, r) : (double)erep;
    }
    case K_SPH_WCA: case K_SPH_TAYLOR:
        return (double)sphere_pair(P, (R)sqrt(r2));
    case K_SCASPA: case K_PSCSPA: case K_CPSCSPA:
        return (double)spsc(P, a, b, r);
    default:
        return 0.0;
    }
}

// Minimum image from fixed-point fractions: the int32 difference of two
// uint32 fractions is the wrapped fraction in [-1/2, 1/2), so no rounding to
// the nearest image is needed (Cuboid::image, geometry.h:110-128).
template <typename R>
SC_HD V3<R> fx_image(uint32_t ax, uint32_t ay, uint32_t az, uint32_t bx, uint32_t by, uint32_t bz, R Lx, R Ly, R Lz) {
    const R s = R(2.3283064365386962890625e-10);  // 2^-32
    return mk<R>(Lx * ((R)(int32_t)(ax - bx) * s), Ly * ((R)(int32_t)(ay - by) * s), Lz * ((R)(int32_t)(az - bz) * s));
}

// FP32 minimum image with the box edge split into a head and a tail
// (L = Lh + Ll exactly to ~2^-48): the rounding of L to FP32 would otherwise
// shift every distance by the same relative amount, a bias that accumulates
// in whole-system sums.
SC_HD V3<float> fx_image_c(uint32_t ax, uint32_t ay, uint32_t az, uint32_t bx, uint32_t by, uint32_t bz,
                           const float* Lh, const float* Ll) {
    const float s = 2.3283064365386962890625e-10f;  // 2^-32
    float x = (float)(int32_t)(ax - bx) * s, y = (float)(int32_t)(ay - by) * s, z = (float)(int32_t)(az - bz) * s;
    return mk<float>(fmaf(x, Lh[0], x * Ll[0]), fmaf(y, Lh[1], y * Ll[1]), fmaf(z, Lh[2], z * Ll[2]));
}

}  // namespace scg
