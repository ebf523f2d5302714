// Triangular-dislocation (TD) elastic field for the traction Green's functions.
//
// Formulation: the TD *displacement* of Nikkhoo & Walter (2015, GJI 201:1119)
// -- full-space angular dislocations, the TD Burgers function, and the
// half-space free-surface correction -- evaluated on forward-mode dual numbers
// seeded with the receiver's Cartesian coordinates.  Every evaluation therefore
// yields the exact displacement gradient G[i][j] = du_i/dX_j; strain, stress and
// fault tractions follow from G.  This replaces the reference's closed-form
// strain expressions (src/TDstressFS_C.cpp:51-121, :613-1486) with the much
// shorter displacement expressions (src/DH_greenfunction.py:71-104, :273-351):
// one reciprocal per distinct denominator instead of one IEEE divide per
// occurrence, and one pass serves every Burgers vector (strike and dip slip).
//
// Configuration logic follows the reference's stress path (the C++ behind the
// ctypes boundary): trimode (TDstressFS_C.cpp:125-166), configuration +1/-1 side
// orientation (:428-438), FSC configuration I/II and the 1e-4 degenerate-side
// threshold (:1514-1552, :1620-1633), trimode-0 zeroing (:446-477), and the
// image sign flip of a TD lying in the free surface (:1998-2007).
//
// Everything on the per-pair path is templated on the real type R so that the
// same source compiles for the GPU (R = double) and for a host flop counter.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PQ_HD __host__ __device__ __forceinline__
#else
#define PQ_HD inline
#endif

#include <cmath>

namespace pq3d {

// Scalar primitives, overloadable by a counting type.
PQ_HD double rsqrt_(double a) { return sqrt(a); }
PQ_HD double rlog_(double a) { return log(a); }
PQ_HD double ratan_(double a) { return atan(a); }
PQ_HD double ratan2_(double a, double b) { return atan2(a, b); }
// 1 / a.  On the device: the hardware estimate refined by two Newton steps
// (within an ulp of the quotient, without the IEEE division's scaling and
// fix-up).  At a = 0 or a = inf it gives NaN where the division gives inf or
// 0; only singular receivers (on a dislocation line or vertex) reach those,
// and their outputs are not finite either way.
PQ_HD double rrcp_(double a) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(a);
    double e = fma(-a, r, 1.0);
    r = fma(r, e, r);
    e = fma(-a, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / a;
#endif
}

// s = sqrt(a) and h = 0.5 / s together.  Generic form: the primitives above
// (so a counting type counts a square root and a reciprocal).  On the device
// for double: the hardware reciprocal square root refined by two coupled
// Newton (Goldschmidt) steps for g -> sqrt(a) and h -> 1 / (2 sqrt(a)), then
// one residual correction of g; a > 0 (a = 0 only at a singular receiver).
template <class R>
PQ_HD void sqrt_hrcp(const R& a, R& s, R& h) {
    s = rsqrt_(a);
    h = R(0.5) * rrcp_(s);
}
PQ_HD void sqrt_hrcp(const double& a, double& s, double& h) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, hh = 0.5 * y;
    double r = fma(-g, hh, 0.5);
    g = fma(g, r, g);
    hh = fma(hh, r, hh);
    r = fma(-g, hh, 0.5);
    g = fma(g, r, g);
    hh = fma(hh, r, hh);
    const double d = fma(-g, g, a);
    s = fma(d, hh, g);
    h = hh;
#else
    s = sqrt(a);
    h = 0.5 / s;
#endif
}

// --------------------------------------------------------------------------
// Forward-mode dual number: value and gradient with respect to (X, Y, Z)
// --------------------------------------------------------------------------
template <class R>
struct Dual {
    R v, x, y, z;
};

template <class R> PQ_HD Dual<R> mkd(R v, R gx, R gy, R gz) { return Dual<R>{v, gx, gy, gz}; }
template <class R> PQ_HD Dual<R> operator+(const Dual<R>& a, const Dual<R>& b) { return {a.v + b.v, a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> PQ_HD Dual<R> operator-(const Dual<R>& a, const Dual<R>& b) { return {a.v - b.v, a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> PQ_HD Dual<R> operator-(const Dual<R>& a) { return {-a.v, -a.x, -a.y, -a.z}; }
template <class R> PQ_HD Dual<R> operator+(const Dual<R>& a, R c) { return {a.v + c, a.x, a.y, a.z}; }
template <class R> PQ_HD Dual<R> operator+(R c, const Dual<R>& a) { return {a.v + c, a.x, a.y, a.z}; }
template <class R> PQ_HD Dual<R> operator-(const Dual<R>& a, R c) { return {a.v - c, a.x, a.y, a.z}; }
template <class R> PQ_HD Dual<R> operator-(R c, const Dual<R>& a) { return {c - a.v, -a.x, -a.y, -a.z}; }
template <class R> PQ_HD Dual<R> operator*(const Dual<R>& a, R c) { return {a.v * c, a.x * c, a.y * c, a.z * c}; }
template <class R> PQ_HD Dual<R> operator*(R c, const Dual<R>& a) { return {a.v * c, a.x * c, a.y * c, a.z * c}; }
template <class R> PQ_HD Dual<R> operator*(const Dual<R>& a, const Dual<R>& b) {
    return {a.v * b.v, a.v * b.x + a.x * b.v, a.v * b.y + a.y * b.v, a.v * b.z + a.z * b.v};
}
template <class R> PQ_HD Dual<R> rcp(const Dual<R>& a) {
    R q = rrcp_(a.v);
    R m = -(q * q);
    return {q, a.x * m, a.y * m, a.z * m};
}
template <class R> PQ_HD Dual<R> dsqrt(const Dual<R>& a) {
    R s, h;
    sqrt_hrcp(a.v, s, h);
    return {s, a.x * h, a.y * h, a.z * h};
}
template <class R> PQ_HD Dual<R> dlog(const Dual<R>& a) {
    R q = rrcp_(a.v);
    return {rlog_(a.v), a.x * q, a.y * q, a.z * q};
}
template <class R> PQ_HD Dual<R> datan(const Dual<R>& a) {
    R q = rrcp_(R(1) + a.v * a.v);
    return {ratan_(a.v), a.x * q, a.y * q, a.z * q};
}
template <class R> PQ_HD Dual<R> datan2(const Dual<R>& n, const Dual<R>& d) {
    R q = rrcp_(n.v * n.v + d.v * d.v);
    R a = d.v * q, b = -(n.v * q);
    return {ratan2_(n.v, d.v), n.x * a + d.x * b, n.y * a + d.y * b, n.z * a + d.z * b};
}

// --------------------------------------------------------------------------
// Per-source geometry, precomputed once per TD (src_setup below)
// --------------------------------------------------------------------------
struct TDFrame {               // one TD (the source, or its image)
    double P2[3];              // TDCS origin
    double ax[3][3];           // rows: Vnorm, Vstrike, Vdip (EFCS components)
    double p1[3], p3[3];       // vertices 1 and 3 in TDCS (vertex 2 is the origin)
    double side[3][2];         // (y, z) of -e13, e12, e23: configuration +1 sides
    double ca[3], sa[3];       // cos, sin of (angle_k - pi) for vertices 1..3
    double bary_den;           // barycentric denominator of trimodefinder
};

struct FSCSide {               // one side PA->PB of the free-surface correction
    double PA[3];
    double e1x, e1y;           // ey1 = (e1x, e1y, 0); ey2 = (e1y, -e1x, 0); ey3 = -eZ
    double yAB[3];             // side vector in the side frame
    double beta;
    double aA, aB;             // -PA.z, -PB.z
    double cs[2][4];           // [configuration I, II]: sinB, cosB, cotB, 1/tan(B/2)
    int skip;                  // |beta| or |pi - beta| below 1e-4: no contribution
    int pad;
};

template <int NB>
struct SrcGeom {
    TDFrame main, image;
    FSCSide side[3];
    double bT[NB][3];          // TDCS Burgers vectors (bx = Ts, by = Ss, bz = Ds)
    double bS[3][NB][3];       // per FSC side: EFCS Burgers vector in the side frame
    int flush;                 // all three vertices at z = 0
    // Sides shared with the neighbouring table entries (-1: none): the side
    // shared with the previous entry, whether the previous entry traverses it
    // the other way, and the side shared with the next entry.
    int prev_side, prev_rev, next_side;
};

PQ_HD void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
PQ_HD double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// TDCS axes (TDstressFS_C.cpp:300-338)
PQ_HD void frame_axes(const double* P1, const double* P2, const double* P3, double ax[3][3]) {
    double u[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    double w[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    double n[3];
    cross3(u, w, n);
    double nn = sqrt(dot3(n, n));
    for (int k = 0; k < 3; ++k) ax[0][k] = n[k] / nn;
    // strike = eZ x n; a horizontal TD takes strike along y (TDstressFS_C.cpp:1826-1829)
    double s[3] = {-ax[0][1], ax[0][0], 0.0};
    double ns = sqrt(dot3(s, s));
    if (ns == 0.0) {
        s[0] = 0.0;
        s[1] = ax[0][2];
        s[2] = 0.0;
    } else {
        s[0] /= ns;
        s[1] /= ns;
    }
    for (int k = 0; k < 3; ++k) ax[1][k] = s[k];
    cross3(ax[0], ax[1], ax[2]);
}

// Vertices, side directions and angles in TDCS (TDstressFS_C.cpp:341-392)
PQ_HD void td_frame_setup(const double* P1, const double* P2, const double* P3, TDFrame& f) {
    const double PI_ = 3.141592653589793;
    frame_axes(P1, P2, P3, f.ax);
    for (int k = 0; k < 3; ++k) f.P2[k] = P2[k];
    double d1[3] = {P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2]};
    double d3[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
    for (int k = 0; k < 3; ++k) {
        f.p1[k] = dot3(f.ax[k], d1);
        f.p3[k] = dot3(f.ax[k], d3);
    }
    double e12[3], e13[3], e23[3];
    for (int k = 0; k < 3; ++k) {
        e12[k] = -f.p1[k];
        e13[k] = f.p3[k] - f.p1[k];
        e23[k] = f.p3[k];
    }
    double n12 = sqrt(dot3(e12, e12)), n13 = sqrt(dot3(e13, e13)), n23 = sqrt(dot3(e23, e23));
    for (int k = 0; k < 3; ++k) {
        e12[k] /= n12;
        e13[k] /= n13;
        e23[k] /= n23;
    }
    double me12[3] = {-e12[0], -e12[1], -e12[2]};
    double ang[3] = {acos(dot3(e12, e13)), acos(dot3(me12, e23)), acos(dot3(e23, e13))};
    f.side[0][0] = -e13[1];
    f.side[0][1] = -e13[2];
    f.side[1][0] = e12[1];
    f.side[1][1] = e12[2];
    f.side[2][0] = e23[1];
    f.side[2][1] = e23[2];
    for (int k = 0; k < 3; ++k) {
        double al = -PI_ + ang[k];
        f.ca[k] = cos(al);
        f.sa[k] = sin(al);
    }
    f.bary_den = (f.p1[1] - f.p3[1]) * (0.0 - f.p3[2]) - (0.0 - f.p3[1]) * (f.p1[2] - f.p3[2]);
}

// One free-surface-correction side (TDstressFS_C.cpp:1504-1633)
PQ_HD void fsc_side_setup(const double* PA, const double* PB, FSCSide& s) {
    const double PI_ = 3.141592653589793;
    double sv[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    double ls = sqrt(dot3(sv, sv));
    double beta = acos(-sv[2] / ls);
    s.beta = beta;
    s.skip = (fabs(beta) < 1.0e-4 || fabs(PI_ - beta) < 1.0e-4) ? 1 : 0;
    s.pad = 0;
    double h = sqrt(sv[0] * sv[0] + sv[1] * sv[1]);
    s.e1x = h != 0.0 ? sv[0] / h : 0.0;
    s.e1y = h != 0.0 ? sv[1] / h : 0.0;
    for (int k = 0; k < 3; ++k) s.PA[k] = PA[k];
    s.yAB[0] = sv[0] * s.e1x + sv[1] * s.e1y;
    s.yAB[1] = sv[0] * s.e1y - sv[1] * s.e1x;
    s.yAB[2] = -sv[2];
    s.aA = -PA[2];
    s.aB = -PB[2];
    double bb[2] = {s.skip ? 0.5 * PI_ : -PI_ + beta, s.skip ? 0.5 * PI_ : beta};
    for (int c = 0; c < 2; ++c) {
        s.cs[c][0] = sin(bb[c]);
        s.cs[c][1] = cos(bb[c]);
        s.cs[c][2] = 1.0 / tan(bb[c]);
        s.cs[c][3] = 1.0 / tan(0.5 * bb[c]);
    }
}

// Complete set-up for one source TD and NB Burgers vectors (TDCS: Ts, Ss, Ds).
template <int NB>
PQ_HD void src_setup(const double* P1, const double* P2, const double* P3, const double (*bT)[3],
                     SrcGeom<NB>& g) {
    td_frame_setup(P1, P2, P3, g.main);
    double Q1[3] = {P1[0], P1[1], -P1[2]}, Q2[3] = {P2[0], P2[1], -P2[2]}, Q3[3] = {P3[0], P3[1], -P3[2]};
    td_frame_setup(Q1, Q2, Q3, g.image);
    fsc_side_setup(P1, P2, g.side[0]);
    fsc_side_setup(P2, P3, g.side[1]);
    fsc_side_setup(P3, P1, g.side[2]);
    g.flush = (P1[2] == 0.0 && P2[2] == 0.0 && P3[2] == 0.0) ? 1 : 0;
    g.prev_side = g.next_side = -1;
    g.prev_rev = 0;
    for (int b = 0; b < NB; ++b) {
        double Bv[3];
        for (int k = 0; k < 3; ++k) {
            g.bT[b][k] = bT[b][k];
            Bv[k] = g.main.ax[0][k] * bT[b][0] + g.main.ax[1][k] * bT[b][1] + g.main.ax[2][k] * bT[b][2];
        }
        for (int sd = 0; sd < 3; ++sd) {
            const FSCSide& s = g.side[sd];
            g.bS[sd][b][0] = Bv[0] * s.e1x + Bv[1] * s.e1y;
            g.bS[sd][b][1] = Bv[0] * s.e1y - Bv[1] * s.e1x;
            g.bS[sd][b][2] = -Bv[2];
        }
    }
}

// --------------------------------------------------------------------------
// Full-space part of one TD (TDdispFS, DH_greenfunction.py:106-208)
// --------------------------------------------------------------------------
// G[b] += grad u for Burgers vector b.  `flip` negates the xz and yz strain of
// this contribution (image of a TD lying in the free surface).
template <class R, int NB, bool BX>
PQ_HD void td_fullspace_grad(const TDFrame& f, R X, R Y, R Z, const double (*bT)[3], R nu, bool flip,
                             R (*G)[9]) {
    const R dX = X - R(f.P2[0]), dY = Y - R(f.P2[1]), dZ = Z - R(f.P2[2]);
    Dual<R> q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        q[k] = mkd<R>(dX * R(f.ax[k][0]) + dY * R(f.ax[k][1]) + dZ * R(f.ax[k][2]), R(f.ax[k][0]),
                      R(f.ax[k][1]), R(f.ax[k][2]));
    // configuration +1 / -1 / 0 (trimodefinder, barycentric in TDCS y-z)
    const R x1 = R(f.p1[1]), y1 = R(f.p1[2]), x3 = R(f.p3[1]), y3 = R(f.p3[2]);
    const R den = R(f.bary_den);
    const R yp = q[1].v, zp = q[2].v;
    const R ba = ((yp - x3) * (R(0) - y3) - (R(0) - x3) * (zp - y3)) / den;
    const R bb = ((x1 - x3) * (zp - y3) - (yp - x3) * (y1 - y3)) / den;
    const R bc = R(1) - ba - bb;
    int tm = 1;
    if ((ba <= R(0) && bb > bc && bc > ba) || (bb <= R(0) && bc > ba && ba > bb) ||
        (bc <= R(0) && ba > bb && bb > bc))
        tm = -1;
    else if ((ba == R(0) && bb >= R(0) && bc >= R(0)) || (ba >= R(0) && bb == R(0) && bc >= R(0)) ||
             (ba >= R(0) && bb >= R(0) && bc == R(0)))
        tm = 0;
    if (tm == 0 && q[0].v != R(0)) tm = 1;
    if (tm == 0) return;  // the stress path leaves these receivers at zero
    const R sgn = tm == -1 ? R(-1) : R(1);
    const R k2 = R(1) - R(2) * nu;

    // angular-dislocation sums (without the 1/(8 pi (1-nu)) factor), TDCS
    Dual<R> U[NB][3];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c) U[b][c] = mkd<R>(R(0), R(0), R(0), R(0));

#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        const R vy = k == 0 ? R(f.p1[1]) : (k == 2 ? R(f.p3[1]) : R(0));
        const R vz = k == 0 ? R(f.p1[2]) : (k == 2 ? R(f.p3[2]) : R(0));
        const R sy = sgn * R(f.side[k][0]);
        const R sz = sgn * R(f.side[k][1]);
        const R cA = R(f.ca[k]), sA = R(f.sa[k]);
        // angular-dislocation coordinates (TDSetupD, DH_greenfunction.py:40-69)
        const Dual<R> dy = q[1] - vy, dz = q[2] - vz;
        const Dual<R> x = q[0];
        const Dual<R> y = dy * sz - dz * sy;
        Dual<R> z = dy * sy + dz * sz;
        // AngDisDisp (DH_greenfunction.py:71-104)
        const Dual<R> eta = y * cA - z * sA;
        Dual<R> zeta = y * sA + z * cA;
        const Dual<R> r = dsqrt(x * x + y * y + z * z);
        if (zeta.v > r.v) zeta = r;
        if (z.v > r.v) z = r;
        const Dual<R> rq = r - zeta, rz = r - z;
        const Dual<R> ir = rcp(r), iq = rcp(rq), iz = rcp(rz);
        const Dual<R> lq = dlog(rq), lz = dlog(rz);
        const Dual<R> o = ir * iq;  // 1 / (r (r - zeta))
        const Dual<R> p = ir * iz;  // 1 / (r (r - z))
        const Dual<R> L = cA * lq - lz;
        const Dual<R> xx = x * x;
        const Dual<R> cop = cA * o - p;
        // by-basis
        const Dual<R> uY = xx * cop - k2 * L;
        const Dual<R> vY = x * (y * cop - (sA * cA) * iq);
        const Dual<R> wY = x * (z * (cA * o) - (cA * cA) * iq + ir);
        // bz-basis (common factor sinA applied below)
        const Dual<R> uZ = k2 * lq - xx * o;
        const Dual<R> vZ = x * (sA * iq - y * o);
        const Dual<R> wZ = x * (cA * iq - z * o);
        Dual<R> uX, vX, wX;
        if (BX) {
            uX = x * (y * p - eta * o);
            vX = eta * (sA * iq - y * o) + y * y * p + k2 * L;
            wX = eta * (cA * iq - z * o) - y * ir - (k2 * sA) * lq;
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const R by = R(bT[b][1]), bz = R(bT[b][2]);
            const R cy = sz * by - sy * bz;
            const R cz = (sy * by + sz * bz) * sA;
            Dual<R> u = cy * uY + cz * uZ;
            Dual<R> v0 = cy * vY + cz * vZ;
            Dual<R> w0 = cy * wY + cz * wZ;
            if (BX) {
                const R bx = R(bT[b][0]);
                u = u + bx * uX;
                v0 = v0 + bx * vX;
                w0 = w0 + bx * wX;
            }
            U[b][0] = U[b][0] + u;
            U[b][1] = U[b][1] + (v0 * sz + w0 * sy);
            U[b][2] = U[b][2] + (w0 * sz - v0 * sy);
        }
    }

    // Burgers function of the whole TD (solid angle, DH_greenfunction.py:182-195)
    Dual<R> a[3], c[3];
    const Dual<R>* bq = q;  // b = p2 - X1 = -q
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        a[k] = R(f.p1[k]) - q[k];
        c[k] = R(f.p3[k]) - q[k];
    }
    const Dual<R> na = dsqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const Dual<R> nb = dsqrt(bq[0] * bq[0] + bq[1] * bq[1] + bq[2] * bq[2]);
    const Dual<R> nc = dsqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    // with b = -q: a.(b x c) = -a.(q x c), a.b = -a.q, b.c = -q.c
    const Dual<R> qxc0 = bq[1] * c[2] - bq[2] * c[1];
    const Dual<R> qxc1 = bq[2] * c[0] - bq[0] * c[2];
    const Dual<R> qxc2 = bq[0] * c[1] - bq[1] * c[0];
    const Dual<R> num = -(a[0] * qxc0 + a[1] * qxc1 + a[2] * qxc2);
    const Dual<R> ab = -(a[0] * bq[0] + a[1] * bq[1] + a[2] * bq[2]);
    const Dual<R> ac = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    const Dual<R> bcd = -(bq[0] * c[0] + bq[1] * c[1] + bq[2] * c[2]);
    const Dual<R> den2 = na * nb * nc + ab * nc + ac * nb + bcd * na;
    const Dual<R> Fi = datan2(num, den2);  // times -1/(2 pi)

    const R K = R(1) / (R(8) * R(3.141592653589793) * (R(1) - nu));
    const R KF = R(-1) / (R(2) * R(3.141592653589793));
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        Dual<R> Ut[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) Ut[k] = K * U[b][k] + (KF * R(bT[b][k])) * Fi;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const R gx = Ut[0].x * R(f.ax[0][i]) + Ut[1].x * R(f.ax[1][i]) + Ut[2].x * R(f.ax[2][i]);
            const R gy = Ut[0].y * R(f.ax[0][i]) + Ut[1].y * R(f.ax[1][i]) + Ut[2].y * R(f.ax[2][i]);
            const R gz = Ut[0].z * R(f.ax[0][i]) + Ut[1].z * R(f.ax[1][i]) + Ut[2].z * R(f.ax[2][i]);
            const R f0 = (flip && i == 2) ? R(-1) : R(1);  // (2,0), (2,1)
            const R f2 = (flip && i != 2) ? R(-1) : R(1);  // (0,2), (1,2)
            G[b][i * 3 + 0] += f0 * gx;
            G[b][i * 3 + 1] += f0 * gy;
            G[b][i * 3 + 2] += f2 * gz;
        }
    }
}

// G[b] += K * w_b * grad(v1 ey1 + v2 ey2 + v3 ey3), w_b = bS[b][k]
template <class R, int NB>
PQ_HD void fsc_scatter(const double (*bS)[3], int k, R K, R e1x, R e1y, const Dual<R>& v1, const Dual<R>& v2,
                       const Dual<R>& v3, R (*G)[9]) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const R w = K * R(bS[b][k]);
        const R a1 = w * e1x, a2 = w * e1y;
        G[b][0] += a1 * v1.x + a2 * v2.x;
        G[b][1] += a1 * v1.y + a2 * v2.y;
        G[b][2] += a1 * v1.z + a2 * v2.z;
        G[b][3] += a2 * v1.x - a1 * v2.x;
        G[b][4] += a2 * v1.y - a1 * v2.y;
        G[b][5] += a2 * v1.z - a1 * v2.z;
        G[b][6] -= w * v3.x;
        G[b][7] -= w * v3.y;
        G[b][8] -= w * v3.z;
    }
}

// G[k] += K * grad(v1 ey1 + v2 ey2 + v3 ey3): the response to the unit
// Burgers vector along side-frame axis k (BASIS mode of the side evaluation)
template <class R>
PQ_HD void fsc_scatter_basis(int k, R K, R e1x, R e1y, const Dual<R>& v1, const Dual<R>& v2, const Dual<R>& v3,
                             R (*G)[9]) {
    const R a1 = K * e1x, a2 = K * e1y;
    G[k][0] += a1 * v1.x + a2 * v2.x;
    G[k][1] += a1 * v1.y + a2 * v2.y;
    G[k][2] += a1 * v1.z + a2 * v2.z;
    G[k][3] += a2 * v1.x - a1 * v2.x;
    G[k][4] += a2 * v1.y - a1 * v2.y;
    G[k][5] += a2 * v1.z - a1 * v2.z;
    G[k][6] -= K * v3.x;
    G[k][7] -= K * v3.y;
    G[k][8] -= K * v3.z;
}

template <class R, int NB, bool BASIS>
PQ_HD void fsc_emit(const double (*bS)[3], int k, R K, R e1x, R e1y, const Dual<R>& v1, const Dual<R>& v2,
                    const Dual<R>& v3, R (*G)[9]) {
    if (BASIS)
        fsc_scatter_basis<R>(k, K, e1x, e1y, v1, v2, v3, G);
    else
        fsc_scatter<R, NB>(bS, k, K, e1x, e1y, v1, v2, v3, G);
}

// --------------------------------------------------------------------------
// Free-surface correction: one end of one side (AngDisDispFSC,
// DH_greenfunction.py:273-351), accumulated as sgn * grad u into G.
// cs = {sinB, cosB, cotB, 1/tan(B/2)} of the configuration in use.
// BASIS: G has three rows, row k the response to the unit Burgers vector
// along side-frame axis k (bS unused); otherwise G[b] is the response to
// Burgers vector bS[b] (side-frame components).
// --------------------------------------------------------------------------
template <class R, int NB, bool BASIS = false>
PQ_HD void fsc_end_grad(const Dual<R>& y1, const Dual<R>& y2, const Dual<R>& y3, const double* cs, R a,
                        R nu, const double (*bS)[3], R e1x, R e1y, R sgn, R (*G)[9]) {
    const R s = R(cs[0]), c = R(cs[1]), ct = R(cs[2]), ith = R(cs[3]);
    const R n = nu, k1 = R(1) - R(2) * nu, k2 = R(1) - nu;
    const Dual<R> y3b = y3 + R(2) * a;
    const Dual<R> z1b = y1 * c + y3b * s;
    const Dual<R> z3b = y3b * c - y1 * s;
    const Dual<R> r2b = y1 * y1 + y2 * y2 + y3b * y3b;
    const Dual<R> rb = dsqrt(r2b);
    const Dual<R> irb = rcp(rb);
    const Dual<R> ir2 = irb * irb;
    const Dual<R> ir3 = ir2 * irb;
    const Dual<R> ry = rb + y3b, rz = rb + z3b;
    const Dual<R> iry = rcp(ry), irz = rcp(rz);
    const Dual<R> lry = dlog(ry), lrz = dlog(rz);
    const Dual<R> W = y3b - a;
    const Dual<R> arb = a * irb;
    const Dual<R> rc = rb * c + y3b;   // rb cosB + y3b
    const Dual<R> rsy = rb * s - y1;   // rb sinB - y1
    const Dual<R> Fib = R(2) * datan(-(y2 * rcp(y1 - ry * ith)));
    const Dual<R> y2W = y2 * W;
    const Dual<R> yy1 = y1 * y1, yy2 = y2 * y2;
    const Dual<R> carb = c + arb;      // cosB + a/rb
    const Dual<R> narb = n + arb;      // nu + a/rb
    const Dual<R> n2arb = R(2) * n + arb;
    const Dual<R> irbiry = irb * iry, irbirz = irb * irz;
    const R ic = R(1) / c;
    const R ct2 = ct * ct;

    // u = v1 ey1 + v2 ey2 + v3 ey3 in EFCS, scaled by sgn / (4 pi (1 - nu)); each
    // basis block goes straight into G so that no dual accumulators stay live.
    const R K = sgn / (R(4) * R(3.141592653589793) * (R(1) - nu));

    // ---- b1 basis
    {
        const Dual<R> v1 =
            (R(-2) * k2 * k1 * ct2) * Fib +
            (k1 * y2 * iry) * ((k1 - arb) * ct - y1 * iry * narb) +
            (k1 * c * ct) * (y2 * irz * carb) +
            (a * ct) * (y2W * ir3) +
            y2W * irbiry * (-k1 * ct + y1 * iry * n2arb + a * y1 * ir2) +
            y2W * irbirz * (c * irz * (rc * (k1 * c - arb) * ct + (R(2) * k2 * c) * rsy) - (a * c * ct) * (y3b * ir2));
        const Dual<R> v2 =
            k1 * ((R(2) * k2 * ct2 - n) * lry - ((R(2) * k2 * ct2 + k1) * c) * lrz) -
            k1 * iry * (y1 * ct * (k1 - arb) + n * y3b - a + yy2 * iry * narb) -
            (k1 * ct) * (z1b * irz * carb) -
            (a * ct) * (y1 * W * ir3) +
            W * iry * (R(-2) * n + irb * (k1 * ct * y1 - a) + yy2 * irbiry * n2arb + a * yy2 * ir3) +
            W * irz * (c * c - irb * ((k1 * ct) * z1b + a * c) + (a * ct) * (y3b * z1b * ir3) -
                       irbirz * ((c * c) * yy2 - (a * ct) * (z1b * irb * rc)));
        const Dual<R> v3 =
            (R(2) * k2) * ((k1 * ct) * Fib + y2 * iry * n2arb - (y2 * c) * irz * carb) +
            y2W * irb * ((R(2) * n) * iry + a * ir2) +
            (y2W * c) * irbirz * (k1 - rc * irz * carb - a * y3b * ir2);
        fsc_emit<R, NB, BASIS>(bS, 0, K, e1x, e1y, v1, v2, v3, G);
    }
    // ---- b2 basis
    {
        const Dual<R> v1 =
            k1 * ((R(2) * k2 * ct2 + n) * lry - ((R(2) * k2 * ct2 + R(1)) * c) * lrz) +
            k1 * iry * ((-k1 * ct) * y1 + n * y3b - a + (a * ct) * (y1 * irb) + yy1 * iry * narb) -
            (k1 * ct) * irz * (c * z1b - (a * ic) * (rsy * irb)) -
            (a * ct) * (y1 * W * ir3) +
            W * iry * (R(2) * n + irb * ((k1 * ct) * y1 + a) - yy1 * irbiry * n2arb - a * yy1 * ir3) +
            (ct * W) * irz * (-(c * s) + (a * ic) * (y1 * y3b * ir3) +
                              rsy * irb * ((R(2) * k2 * c) - rc * irz * (R(1) + arb * ic)));
        const Dual<R> v2 =
            (R(2) * k2 * k1 * ct2) * Fib +
            (k1 * y2 * iry) * (-((k1 - arb) * ct) + y1 * iry * narb) -
            (k1 * ct) * (y2 * irz * (R(1) + arb * ic)) -
            (a * ct) * (y2W * ir3) +
            y2W * irbiry * (k1 * ct - (R(2) * n) * y1 * iry - a * y1 * irb * (irb + iry)) +
            (ct * y2W) * irbirz * (R(-2) * k2 * c + rc * irz * (R(1) + arb * ic) + (a * ic) * (y3b * ir2));
        const Dual<R> v3 =
            (R(-2) * k2 * k1 * ct) * (lry - c * lrz) -
            (R(2) * k2) * (y1 * iry * n2arb) +
            (R(2) * k2) * (z1b * irz * carb) +
            W * irb * (k1 * ct - (R(2) * n) * y1 * iry - a * y1 * ir2) -
            W * irz * (c * s + ct * rc * irb * ((R(2) * k2 * c) - rc * irz) +
                       arb * (s - y3b * z1b * ir2 - z1b * rc * irbirz));
        fsc_emit<R, NB, BASIS>(bS, 1, K, e1x, e1y, v1, v2, v3, G);
    }
    // ---- b3 basis
    {
        const Dual<R> v1 =
            k1 * (y2 * iry * (R(1) + arb) - (y2 * c) * irz * carb) -
            y2W * irb * (a * ir2 + iry) +
            (y2W * c) * irbirz * (rc * irz * carb + a * y3b * ir2);
        const Dual<R> v2 =
            k1 * (-(s * lrz) - y1 * iry * (R(1) + arb) + z1b * irz * carb) +
            y1 * W * irb * (a * ir2 + iry) -
            W * irz * (s * (c - arb) + z1b * irb * (R(1) + a * y3b * ir2) -
                       irbirz * ((c * s) * yy2 - a * z1b * irb * rc));
        const Dual<R> v3 =
            (R(2) * k2) * Fib + (R(2) * k2 * s) * (y2 * irz * carb) +
            (y2W * s) * irbirz * (R(1) + rc * irz * carb + a * y3b * ir2);
        fsc_emit<R, NB, BASIS>(bS, 2, K, e1x, e1y, v1, v2, v3, G);
    }
}

// One side of the free-surface correction (AngSetupFSC, DH_greenfunction.py:234-271)
// BASIS: G's three rows take the responses to the unit Burgers vectors along
// the side-frame axes (bS unused), as in fsc_end_grad.
template <class R, int NB, bool BASIS = false>
PQ_HD void fsc_side_grad(const FSCSide& sd, const double (*bS)[3], R X, R Y, R Z, R nu, R (*G)[9]) {
    if (sd.skip) return;
    const R e1x = R(sd.e1x), e1y = R(sd.e1y);
    const R dX = X - R(sd.PA[0]), dY = Y - R(sd.PA[1]), dZ = Z - R(sd.PA[2]);
    const Dual<R> yA1 = mkd<R>(dX * e1x + dY * e1y, e1x, e1y, R(0));
    const Dual<R> yA2 = mkd<R>(dX * e1y - dY * e1x, e1y, -e1x, R(0));
    const Dual<R> yA3 = mkd<R>(-dZ, R(0), R(0), R(-1));
    const int cfg = (R(sd.beta) * yA1.v >= R(0)) ? 0 : 1;  // configuration I : II
    const double* cs = sd.cs[cfg];
    // end A enters with -1, end B (shifted by the side vector) with +1
#pragma unroll 1
    for (int e = 0; e < 2; ++e) {
        const R o0 = e ? R(sd.yAB[0]) : R(0), o1 = e ? R(sd.yAB[1]) : R(0), o2 = e ? R(sd.yAB[2]) : R(0);
        fsc_end_grad<R, NB, BASIS>(yA1 - o0, yA2 - o1, yA3 - o2, cs, e ? R(sd.aB) : R(sd.aA), nu, bS, e1x, e1y,
                            e ? R(1) : R(-1), G);
    }
}

// --------------------------------------------------------------------------
// Whole pair: G[b] = displacement gradient at (X, Y, Z) for Burgers vector b
// (TDstressHS = main + FSC + image, TDstressFS_C.cpp:1934-2033)
// --------------------------------------------------------------------------
template <class R, int NB, bool BX>
PQ_HD void td_pair_grad(const SrcGeom<NB>& g, R X, R Y, R Z, R nu, bool halfspace, R (*G)[9]) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int k = 0; k < 9; ++k) G[b][k] = R(0);
    // main TD, then (half space) its image in z = 0
#pragma unroll 1
    for (int t = 0; t < (halfspace ? 2 : 1); ++t)
        td_fullspace_grad<R, NB, BX>(t ? g.image : g.main, X, Y, Z, g.bT, nu, t == 1 && g.flush != 0, G);
    if (halfspace) {
#pragma unroll 1
        for (int s = 0; s < 3; ++s) fsc_side_grad<R, NB>(g.side[s], g.bS[s], X, Y, Z, nu, G);
    }
}

// --------------------------------------------------------------------------
// Sides shared by consecutive sources.  The free-surface correction of a side
// is linear in the Burgers vector, and that of side BA is minus that of side
// AB (the angular-dislocation pair traversed the other way); in BA's frame
// the axes ey1, ey2 are reversed and ey3 kept.  So the responses of side AB
// to the three unit Burgers vectors of its frame, R[k], serve both sources:
//   AB's source:  G[b] += sum_k bS_AB[b][k] R[k]
//   BA's source:  G[b] += bS_BA[b][0] R[0] + bS_BA[b][1] R[1] - bS_BA[b][2] R[2]
// (a neighbour that runs the side the same way uses its bS unchanged).
// --------------------------------------------------------------------------

// The side of triangle A (side k runs from vertex k to vertex k + 1 mod 3, as
// the correction's sides are set up) that triangle B also has, when the two
// share exactly two vertices (bitwise-equal coordinates); -1 otherwise.
// rev: B runs that side the other way.
PQ_HD bool same_point(const double* a, const double* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }
PQ_HD int shared_side(const double (*A)[3], const double (*B)[3], int* rev) {
    int match[3] = {-1, -1, -1};  // B's vertex at each of A's vertices
    int c = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            if (same_point(A[a], B[b])) {
                if (match[a] < 0) match[a] = b;
                ++c;
            }
    *rev = 0;
    if (c != 2) return -1;
    for (int k = 0; k < 3; ++k) {
        const int b0 = match[k], b1 = match[(k + 1) % 3];
        if (b0 < 0 || b1 < 0 || b0 == b1) continue;
        *rev = (b1 == (b0 + 1) % 3) ? 0 : 1;
        return k;
    }
    return -1;
}

// One pair of a run of sources: as td_pair_grad, but the side g.next_side
// (shared with the next source) is evaluated for the three unit Burgers
// vectors and handed on through store(k, c, value), and when use_prev is set
// the side g.prev_side takes the previous source's responses through
// load(k, c) instead of being evaluated.  The responses are stored before the
// previous ones are loaded, so a caller can keep two alternating buffers.
template <class R, class Store, class Load>
PQ_HD void td_pair_grad_shared(const SrcGeom<2>& g, R X, R Y, R Z, R nu, bool halfspace, bool use_prev,
                               Store store, Load load, R (*G)[9]) {
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int k = 0; k < 9; ++k) G[b][k] = R(0);
    int ns = -1, ps = -1;
    if (halfspace) {
        ns = g.next_side;
        ps = use_prev ? g.prev_side : -1;
        if (ns >= 0) {
            R Rb[3][9];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int c = 0; c < 9; ++c) Rb[k][c] = R(0);
            fsc_side_grad<R, 3, true>(g.side[ns], nullptr, X, Y, Z, nu, Rb);
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int c = 0; c < 9; ++c) store(k, c, Rb[k][c]);
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const R w = R(g.bS[ns][b][k]);
#pragma unroll
                    for (int c = 0; c < 9; ++c) G[b][c] += w * Rb[k][c];
                }
        }
        if (ps >= 0) {
            const double s2 = g.prev_rev ? -1.0 : 1.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int c = 0; c < 9; ++c) {
                    const R v = load(k, c);
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        const R w = R(k == 2 ? s2 * g.bS[ps][b][2] : g.bS[ps][b][k]);
                        G[b][c] += w * v;
                    }
                }
        }
    }
    // main TD, then (half space) its image in z = 0
#pragma unroll 1
    for (int t = 0; t < (halfspace ? 2 : 1); ++t)
        td_fullspace_grad<R, 2, false>(t ? g.image : g.main, X, Y, Z, g.bT, nu, t == 1 && g.flush != 0, G);
    if (halfspace) {
#pragma unroll 1
        for (int sd = 0; sd < 3; ++sd)
            if (sd != ns && sd != ps) fsc_side_grad<R, 2>(g.side[sd], g.bS[sd], X, Y, Z, nu, G);
    }
}

// Strain (symmetric gradient) and stress (Hooke), components [xx, yy, zz, xy, xz, yz]
template <class R>
PQ_HD void grad_to_strain_stress(const R* G, R mu, R lam, R* eps, R* sig) {
    eps[0] = G[0];
    eps[1] = G[4];
    eps[2] = G[8];
    eps[3] = R(0.5) * (G[1] + G[3]);
    eps[4] = R(0.5) * (G[2] + G[6]);
    eps[5] = R(0.5) * (G[5] + G[7]);
    const R tr = eps[0] + eps[1] + eps[2];
    const R m2 = R(2) * mu;
    sig[0] = m2 * eps[0] + lam * tr;
    sig[1] = m2 * eps[1] + lam * tr;
    sig[2] = m2 * eps[2] + lam * tr;
    sig[3] = m2 * eps[3];
    sig[4] = m2 * eps[4];
    sig[5] = m2 * eps[5];
}

// Fault tractions on a receiver with frame rows e1, e2, n (GetTtstress,
// QDsim_gpu.py:1400-1417): t = sigma n; returns (t.e1, t.e2, t.n)
template <class R>
PQ_HD void stress_to_traction(const R* s, const R* ev, R* out) {
    const R n0 = ev[6], n1 = ev[7], n2 = ev[8];
    const R t0 = s[0] * n0 + s[3] * n1 + s[4] * n2;
    const R t1 = s[3] * n0 + s[1] * n1 + s[5] * n2;
    const R t2 = s[4] * n0 + s[5] * n1 + s[2] * n2;
    out[0] = t0 * ev[0] + t1 * ev[1] + t2 * ev[2];
    out[1] = t0 * ev[3] + t1 * ev[4] + t2 * ev[5];
    out[2] = t0 * n0 + t1 * n1 + t2 * n2;
}

}  // namespace pq3d
