// sweep_core.h -- device core of the spherical FD eikonal sweep (FDtomoC
// src/sphfd.c), shared by the gfx950 kernel and by host-side builds used in
// tests.  Everything here is header-only and free of HIP runtime calls.
//
// Contents
//   * the four stencils (fdsph3d sphfd.c:5906-6013, fdsphne 6016-6074,
//     fdsph2d 6077-6143, fdsphnf 6146-6184), FP64, same expression trees;
//   * per-face trial specifications: for each of the six faces the 22 trials
//     in the reference's order, each stencil point as an (dx,dy,dz) offset in
//     the slot order the reference loads it (that order fixes the FP64
//     summation order), gates, grid bounds, box-interior conditions, distance
//     and 1D-step expressions;
//   * eval_node<FACE>(): one node's trial sequence and store decision, driven
//     through an accessor so the caller decides which time value (final
//     or pre-face) each in-face neighbour contributes.
//
// Must be compiled with FP contraction off (-ffp-contract=off) so that every
// multiply and add rounds separately, as the reference's x86-64 build does.
#pragma once

#include <stdint.h>

#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FD_HD __host__ __device__
#define FD_INLINE __host__ __device__ __forceinline__
#define FD_NOINLINE __host__ __device__ __attribute__((noinline))
#else
#include <math.h>
#define FD_HD
#define FD_INLINE inline __attribute__((always_inline))
#define FD_NOINLINE __attribute__((noinline))
#endif

namespace fdsweep {

// Face identifiers in the reference's sweep order (sphfd.c:1021, 1770, 2506,
// 3275, 4035, 4800).
enum Face : int { TOP = 0, BOTTOM = 1, EAST = 2, WEST = 3, NORTH = 4, SOUTH = 5 };
// Head-wave counter slot per face (sphfd.c:1759, 2495, 3264, 4024, 4789, 5548).
FD_INLINE int face_slot(int f) {
    return f == TOP ? 5 : f == BOTTOM ? 6 : f == EAST ? 2 : f == WEST ? 1 : f == NORTH ? 3 : 4;
}

constexpr double kBig = 1.0e9;      // "computed" threshold used by every gate
constexpr double kAcausal = 1.e11;  // stencil rejection value
constexpr double kHeadTest = 1.e-3; // sphfd.c:360

// Calls fn(integral_constant<int, I>) for I = I0 .. N-1.
template <int I, int N, class Fn>
FD_INLINE void static_for(Fn &&fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

// ---------------------------------------------------------------- division

// Correctly rounded n / d from y = RN(1/d) (Markstein: with y the correctly
// rounded reciprocal and q0 = RN(n*y) within one ulp of n/d, the residual
// d*q0 - n is exact and q0 - residual*y rounds to the correctly rounded
// quotient, barring overflow and underflow).  Three FP64 operations instead
// of a full division.  Special operands (infinite or NaN numerator, zero or
// infinite divisor) get the IEEE result: on the GPU from v_div_fixup_f64, the
// special-case step of the hardware division sequence, which otherwise
// passes the quotient through; on the host from the division itself.  Every
// divisor passed here is a grid distance or a geometric constant with a
// precomputed reciprocal; the operand magnitudes of the stencils (float
// times and slownesses, kilometre distances) keep the residual far from
// underflow, so results are bit-identical to `n / d`.
FD_INLINE double xdiv(double n, double d, double y) {
    const double q0 = n * y;
    const double e = fma(d, q0, -n);
    const double q = fma(-e, y, q0);
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_div_fixup(q, d, n);
#else
    return __builtin_isfinite(q) && __builtin_isfinite(d) && d != 0.0 ? q : n / d;
#endif
}

// xdiv for a numerator and divisor that are finite by construction (sums of
// precomputed geometric reciprocals over products of two layer radii): the
// special-case step is not needed.  A zero numerator here is +0 (a sum of
// opposite values rounds to +0), whose quotient the formula gives as +0.
FD_INLINE double xdiv_geo(double n, double d, double y) {
    const double q0 = n * y;
    const double e = fma(d, q0, -n);
    return fma(-e, y, q0);
}

// ---------------------------------------------------------------- stencils
//
// The stencils of sphfd.c with every sub-expression that depends on the grid
// geometry only (not on times or slownesses) taken from a per-table geometry
// table (below), where it is computed with the same IEEE operations in the
// same order; divisions by grid distances use xdiv.  The remaining arithmetic
// follows the reference's expression trees operation by operation.

// A grid distance with its correctly rounded reciprocal.
struct Dist {
    double d, y;
};

// New-edge stencil (fdsphne, sphfd.c:6016-6074): points 0..4 known, point 5
// the node; a = 1/d25^2 + 1/d45^2 and ya = RN(1/(2a)) from the table.
FD_INLINE double st_ne(const double t[8], const double s[8], Dist d01, Dist d12, Dist d14, Dist d25, Dist d45,
                       double a, double ya) {
    const double slo = .25 * (s[1] + s[2] + s[4] + s[5]);
    const double dt1 = xdiv(t[4] - t[1], d14.d, d14.y) - xdiv(t[2], d25.d, d25.y);
    const double dt2 = xdiv(t[2] - t[1], d12.d, d12.y) - xdiv(t[4], d45.d, d45.y);
    const double dt3 = xdiv(t[0] - t[3], d01.d, d01.y);
    const double b = 2 * (xdiv(dt1, d25.d, d25.y) + xdiv(dt2, d45.d, d45.y));
    const double c = dt1 * dt1 + dt2 * dt2 + dt3 * dt3 - 4 * slo * slo;
    const double rad = b * b - 4.0 * a * c;
    if (!(rad >= 0.)) return kAcausal;
    const double sr = sqrt(rad);
    const double a2 = 2 * a;
    const double x1 = xdiv(-b + sr, a2, ya);
    const double x2 = xdiv(-b - sr, a2, ya);
    bool bad1 = false, bad2 = false;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        bad1 |= (x1 < t[i]);
        bad2 |= (x2 < t[i]);
    }
    if (bad1) return bad2 ? kAcausal : x2;
    if (bad2) return x1;
    return (x2 < x1) ? x2 : x1;
}

// 2D face stencil (fdsph2d, sphfd.c:6077-6143).  Causality is tested against
// t[0] and t[4] as well, which the caller carries over from earlier trials
// (the reference's 2D call sites never set them); a = 1/d01^2 + 1/d02^2 and
// ya = RN(1/(2a)) from the table.
FD_INLINE double st_2d(const double t[8], const double s[8], Dist d01, Dist d02, Dist d13, Dist d23, double a,
                       double ya) {
    const double slo = .25 * (s[0] + s[1] + s[2] + s[3]);
    const double t1mt3 = t[1] - t[3], t2mt3 = t[2] - t[3];
    const double t1t3t2 = xdiv(t1mt3, d13.d, d13.y) - xdiv(t[2], d02.d, d02.y);
    const double t2t3t1 = xdiv(t2mt3, d23.d, d23.y) - xdiv(t[1], d01.d, d01.y);
    const double b = 2.0 * (xdiv(t1t3t2, d02.d, d02.y) + xdiv(t2t3t1, d01.d, d01.y));
    const double c = t1t3t2 * t1t3t2 + t2t3t1 * t2t3t1 - 4 * slo * slo;
    const double rad = b * b - 4.0 * a * c;
    if (!(rad >= 0.)) return kAcausal;
    const double sr = sqrt(rad);
    const double a2 = 2 * a;
    const double x1 = xdiv(-b + sr, a2, ya);
    const double x2 = xdiv(-b - sr, a2, ya);
    const bool bad1 = (x1 < t[1]) | (x1 < t[2]) | (x1 < t[3]) | (x1 < t[4]) | (x1 < t[0]);
    const bool bad2 = (x2 < t[1]) | (x2 < t[2]) | (x2 < t[3]) | (x2 < t[4]) | (x2 < t[0]);
    if (bad1) return bad2 ? kAcausal : x2;
    if (bad2) return x1;
    return (x2 < x1) ? x2 : x1;
}

// New-face stencil (fdsphnf, sphfd.c:6146-6184): explicit step from the inner layer.
FD_INLINE double st_nf(const double t[8], const double s[8], Dist d02, Dist d12, double d25) {
    const double slo = .5 * (s[5] + s[2]);
    const double t4t0 = xdiv(t[4] - t[0], d02.d, d02.y);
    const double t1t3 = xdiv(t[1] - t[3], d12.d, d12.y);
    double x = slo * slo - 0.25 * (t4t0 * t4t0 + t1t3 * t1t3);
    if (x >= 0.) {
        x = t[2] + d25 * sqrt(x);
        if ((x < t[0]) || (x < t[1]) || (x < t[3]) || (x < t[4])) x = kAcausal;
    } else {
        x = kAcausal;
    }
    return x;
}

// ------------------------------------------------------------ face specs

enum Kind : uint8_t { K3D, KNE, K2D, KNF, K1D };
enum Gate : uint8_t { GLT = 0, GNZ = 1 };               // value < 1e9 | value != 0
enum DistForm : uint8_t { DNONE, DH, DRDQ, DRSDF };      // h | rc*dq | (rc*sin q)*df
enum Factor : uint8_t { XNONE, XH, XR, XDQ, XDF, XS };   // 1D step factors, left to right
enum Axis : int8_t { AXN = -1, AXX = 0, AXY = 1, AXZ = 2 };

struct P3 { int8_t dx, dy, dz; };
struct DS { uint8_t form; int8_t dz, dy; };
struct Trial {
    uint8_t kind;
    int8_t bax0, bdir0, bax1, bdir1;   // grid bounds: dir>0 -> c < n-1, dir<0 -> c > 0
    int8_t iax0, iax1;                 // strict box-interior axes
    uint8_t ngate;
    P3 gate[3];
    uint8_t gtest[3];
    P3 pt[7];
    DS d[5];
    uint8_t fhead;
    DS fac[3];
};

#define FD_P(x, y, z) P3{x, y, z}
#define FD_Z P3{0, 0, 0}
#define FD_D0 DS{DNONE, 0, 0}
#define FD_DH DS{DH, 0, 0}
#define FD_DQ(z) DS{DRDQ, z, 0}
#define FD_DS(z, y) DS{DRSDF, z, y}
#define FD_XH DS{XH, 0, 0}
#define FD_XR(z) DS{XR, z, 0}
#define FD_XQ DS{XDQ, 0, 0}
#define FD_XF DS{XDF, 0, 0}
#define FD_XS(y) DS{XS, 0, y}

// 3D corner: grid bounds on two axes, three in-plane gate points (gate3 carries
// the quirky truthiness test of EAST's first corner), seven points.
#define FD_T3(a0, d0, a1, d1, g0, g1, g2, gt2, p0, p1, p2, p3, p4, p5, p6)                                        \
    Trial{K3D, a0, d0, a1, d1, AXN, AXN, 3, {g0, g1, g2}, {GLT, GLT, gt2}, {p0, p1, p2, p3, p4, p5, p6},        \
          {FD_D0, FD_D0, FD_D0, FD_D0, FD_D0}, 0, {FD_D0, FD_D0, FD_D0}}
#define FD_TNE(a0, d0, ia, g0, gt, p0, p1, p2, p3, p4, d01, d12, d14, d25, d45)                                  \
    Trial{KNE, a0, d0, AXN, 0, ia, AXN, 1, {g0, FD_Z, FD_Z}, {gt, GLT, GLT}, {p0, p1, p2, p3, p4, FD_Z, FD_Z},  \
          {d01, d12, d14, d25, d45}, 0, {FD_D0, FD_D0, FD_D0}}
#define FD_T2A(a0, d0, g0, p1, p2, p3, d01, d02, d13, d23)                                                        \
    Trial{K2D, a0, d0, AXN, 0, AXN, AXN, 1, {g0, FD_Z, FD_Z}, {GLT, GLT, GLT}, {p1, p2, p3, FD_Z, FD_Z, FD_Z, FD_Z}, \
          {d01, d02, d13, d23, FD_D0}, 0, {FD_D0, FD_D0, FD_D0}}
#define FD_T2B(a0, d0, a1, d1, g0, g1, g2, p1, p2, p3, d01, d02, d13, d23)                                        \
    Trial{K2D, a0, d0, a1, d1, AXN, AXN, 3, {g0, g1, g2}, {GLT, GLT, GLT}, {p1, p2, p3, FD_Z, FD_Z, FD_Z, FD_Z},  \
          {d01, d02, d13, d23, FD_D0}, 1, {FD_D0, FD_D0, FD_D0}}
#define FD_TNF(ia0, ia1, p0, p1, p2, p3, p4, d02, d12, d25)                                                       \
    Trial{KNF, AXN, 0, AXN, 0, ia0, ia1, 0, {FD_Z, FD_Z, FD_Z}, {GLT, GLT, GLT}, {p0, p1, p2, p3, p4, FD_Z, FD_Z}, \
          {d02, d12, d25, FD_D0, FD_D0}, 0, {FD_D0, FD_D0, FD_D0}}
#define FD_T1P(nb, f0, f1, f2)                                                                                    \
    Trial{K1D, AXN, 0, AXN, 0, AXN, AXN, 0, {FD_Z, FD_Z, FD_Z}, {GLT, GLT, GLT}, {nb, FD_Z, FD_Z, FD_Z, FD_Z, FD_Z, FD_Z}, \
          {FD_D0, FD_D0, FD_D0, FD_D0, FD_D0}, 0, {f0, f1, f2}}
#define FD_T1(a0, d0, nb, f0, f1, f2)                                                                             \
    Trial{K1D, a0, d0, AXN, 0, AXN, AXN, 1, {nb, FD_Z, FD_Z}, {GLT, GLT, GLT}, {nb, FD_Z, FD_Z, FD_Z, FD_Z, FD_Z, FD_Z}, \
          {FD_D0, FD_D0, FD_D0, FD_D0, FD_D0}, 1, {f0, f1, f2}}

template <int F>
struct FaceSpec;

// TOP: face z1, inner layer z1+1, sphfd.c:1021-1768.
template <>
struct FaceSpec<TOP> {
    static constexpr int nax = AXZ, ndir = +1, a_ax = AXX, b_ax = AXY;
    static constexpr Trial trials[22] = {
        FD_T3(AXY, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 1, 0), FD_P(0, 1, 0), GLT,
              FD_P(1, 1, 1), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(1, 0, 1), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0)),
        FD_T3(AXY, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 1, 0), FD_P(0, 1, 0), GLT,
              FD_P(0, 1, 1), FD_P(-1, 1, 1), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(0, 1, 0), FD_P(-1, 1, 0), FD_P(-1, 0, 0)),
        FD_T3(AXY, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, -1, 0), FD_P(0, -1, 0), GLT,
              FD_P(1, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 1), FD_P(1, -1, 1), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0)),
        FD_T3(AXY, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, -1, 0), FD_P(0, -1, 0), GLT,
              FD_P(0, 0, 1), FD_P(-1, 0, 1), FD_P(-1, -1, 1), FD_P(0, -1, 1), FD_P(-1, 0, 0), FD_P(-1, -1, 0), FD_P(0, -1, 0)),
        FD_TNE(AXX, -1, AXY, FD_P(-1, 0, 0), GLT, FD_P(-1, 1, 1), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(-1, -1, 1),
               FD_P(-1, 0, 0), FD_DQ(1), FD_DS(1, 0), FD_DH, FD_DH, FD_DS(0, 0)),
        FD_TNE(AXX, 1, AXY, FD_P(1, 0, 0), GLT, FD_P(1, -1, 1), FD_P(1, 0, 1), FD_P(0, 0, 1), FD_P(1, 1, 1),
               FD_P(1, 0, 0), FD_DQ(1), FD_DS(1, 0), FD_DH, FD_DH, FD_DS(0, 0)),
        FD_TNE(AXY, 1, AXX, FD_P(0, 1, 0), GLT, FD_P(1, 1, 1), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(-1, 1, 1),
               FD_P(0, 1, 0), FD_DS(1, 1), FD_DQ(1), FD_DH, FD_DH, FD_DQ(0)),
        // sphfd.c:1471 tests this neighbour for truthiness only
        FD_TNE(AXY, -1, AXX, FD_P(0, -1, 0), GNZ, FD_P(-1, -1, 1), FD_P(0, -1, 1), FD_P(0, 0, 1), FD_P(1, -1, 1),
               FD_P(0, -1, 0), FD_DS(1, -1), FD_DQ(1), FD_DH, FD_DH, FD_DQ(0)),
        FD_T2A(AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 0), FD_P(0, 0, 1), FD_P(1, 0, 1), FD_DS(0, 0), FD_DH, FD_DH, FD_DS(1, 0)),
        FD_T2A(AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 0), FD_P(0, 0, 1), FD_P(0, -1, 1), FD_DQ(0), FD_DH, FD_DH, FD_DQ(1)),
        FD_T2A(AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 0), FD_P(0, 0, 1), FD_P(-1, 0, 1), FD_DS(0, 0), FD_DH, FD_DH, FD_DS(1, 0)),
        FD_T2A(AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 0), FD_P(0, 0, 1), FD_P(0, 1, 1), FD_DQ(0), FD_DH, FD_DH, FD_DQ(1)),
        FD_T2B(AXY, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0), FD_P(0, 1, 0), FD_P(1, 1, 0),
               FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, 1)),
        FD_T2B(AXY, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, -1, 0), FD_P(0, -1, 0), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0),
               FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, -1)),
        FD_T2B(AXY, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 1, 0), FD_P(0, 1, 0), FD_P(-1, 0, 0), FD_P(0, 1, 0), FD_P(-1, 1, 0),
               FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, 1)),
        FD_T2B(AXY, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, -1, 0), FD_P(0, -1, 0), FD_P(-1, 0, 0), FD_P(0, -1, 0),
               FD_P(-1, -1, 0), FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, -1)),
        FD_TNF(AXX, AXY, FD_P(0, 1, 1), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(1, 0, 1), FD_P(0, -1, 1), FD_DQ(1), FD_DS(1, 0),
               FD_DH),
        FD_T1P(FD_P(0, 0, 1), FD_XH, FD_D0, FD_D0),
        FD_T1(AXX, 1, FD_P(1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXX, -1, FD_P(-1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXY, 1, FD_P(0, 1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXY, -1, FD_P(0, -1, 0), FD_XR(0), FD_XQ, FD_D0),
    };
};

// BOTTOM: face z2, inner layer z2-1, sphfd.c:1770-2504.
template <>
struct FaceSpec<BOTTOM> {
    static constexpr int nax = AXZ, ndir = -1, a_ax = AXX, b_ax = AXY;
    static constexpr Trial trials[22] = {
        FD_T3(AXY, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 1, 0), FD_P(0, 1, 0), GLT,
              FD_P(1, 1, -1), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(1, 0, -1), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0)),
        FD_T3(AXY, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 1, 0), FD_P(0, 1, 0), GLT,
              FD_P(0, 1, -1), FD_P(-1, 1, -1), FD_P(-1, 0, -1), FD_P(0, 0, -1), FD_P(0, 1, 0), FD_P(-1, 1, 0), FD_P(-1, 0, 0)),
        FD_T3(AXY, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, -1, 0), FD_P(0, -1, 0), GLT,
              FD_P(1, 0, -1), FD_P(0, 0, -1), FD_P(0, -1, -1), FD_P(1, -1, -1), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0)),
        FD_T3(AXY, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, -1, 0), FD_P(0, -1, 0), GLT,
              FD_P(0, 0, -1), FD_P(-1, 0, -1), FD_P(-1, -1, -1), FD_P(0, -1, -1), FD_P(-1, 0, 0), FD_P(-1, -1, 0),
              FD_P(0, -1, 0)),
        FD_TNE(AXX, -1, AXY, FD_P(-1, 0, 0), GLT, FD_P(-1, -1, -1), FD_P(-1, 0, -1), FD_P(0, 0, -1), FD_P(-1, 1, -1),
               FD_P(-1, 0, 0), FD_DQ(-1), FD_DS(-1, 0), FD_DH, FD_DH, FD_DS(0, 0)),
        FD_TNE(AXX, 1, AXY, FD_P(1, 0, 0), GLT, FD_P(1, 1, -1), FD_P(1, 0, -1), FD_P(0, 0, -1), FD_P(1, -1, -1),
               FD_P(1, 0, 0), FD_DQ(-1), FD_DS(-1, 0), FD_DH, FD_DH, FD_DS(0, 0)),
        FD_TNE(AXY, 1, AXX, FD_P(0, 1, 0), GLT, FD_P(-1, 1, -1), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(1, 1, -1),
               FD_P(0, 1, 0), FD_DS(-1, 1), FD_DQ(-1), FD_DH, FD_DH, FD_DQ(0)),
        FD_TNE(AXY, -1, AXX, FD_P(0, -1, 0), GLT, FD_P(1, -1, -1), FD_P(0, -1, -1), FD_P(0, 0, -1), FD_P(-1, -1, -1),
               FD_P(0, -1, 0), FD_DS(-1, -1), FD_DQ(-1), FD_DH, FD_DH, FD_DQ(0)),
        FD_T2A(AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 0), FD_P(0, 0, -1), FD_P(1, 0, -1), FD_DS(0, 0), FD_DH, FD_DH, FD_DS(-1, 0)),
        FD_T2A(AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 0), FD_P(0, 0, -1), FD_P(0, -1, -1), FD_DQ(0), FD_DH, FD_DH, FD_DQ(-1)),
        FD_T2A(AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 0), FD_P(0, 0, -1), FD_P(-1, 0, -1), FD_DS(0, 0), FD_DH, FD_DH,
               FD_DS(-1, 0)),
        FD_T2A(AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 0), FD_P(0, 0, -1), FD_P(0, 1, -1), FD_DQ(0), FD_DH, FD_DH, FD_DQ(-1)),
        FD_T2B(AXY, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0), FD_P(0, 1, 0), FD_P(1, 1, 0),
               FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, 1)),
        FD_T2B(AXY, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, -1, 0), FD_P(0, -1, 0), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0),
               FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, -1)),
        FD_T2B(AXY, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 1, 0), FD_P(0, 1, 0), FD_P(-1, 0, 0), FD_P(0, 1, 0), FD_P(-1, 1, 0),
               FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, 1)),
        FD_T2B(AXY, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, -1, 0), FD_P(0, -1, 0), FD_P(-1, 0, 0), FD_P(0, -1, 0),
               FD_P(-1, -1, 0), FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, -1)),
        FD_TNF(AXX, AXY, FD_P(0, 1, -1), FD_P(-1, 0, -1), FD_P(0, 0, -1), FD_P(1, 0, -1), FD_P(0, -1, -1), FD_DQ(-1),
               FD_DS(-1, 0), FD_DH),
        FD_T1P(FD_P(0, 0, -1), FD_XH, FD_D0, FD_D0),
        FD_T1(AXX, 1, FD_P(1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXX, -1, FD_P(-1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXY, 1, FD_P(0, 1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXY, -1, FD_P(0, -1, 0), FD_XR(0), FD_XQ, FD_D0),
    };
};

// EAST: face x2, inner layer x2-1, in-face axes j (a) and k (b), sphfd.c:2506-3273.
template <>
struct FaceSpec<EAST> {
    static constexpr int nax = AXX, ndir = -1, a_ax = AXY, b_ax = AXZ;
    static constexpr Trial trials[22] = {
        // sphfd.c:2544 tests the third neighbour for truthiness only
        FD_T3(AXZ, 1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 1), FD_P(0, 0, 1), GNZ,
              FD_P(-1, 1, 1), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(-1, 0, 1), FD_P(-1, 1, 0), FD_P(0, 1, 0), FD_P(-1, 0, 0)),
        FD_T3(AXZ, 1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 1), FD_P(0, 0, 1), GLT,
              FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 1), FD_P(-1, -1, 1), FD_P(-1, 0, 0), FD_P(0, -1, 0), FD_P(-1, -1, 0)),
        FD_T3(AXZ, -1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, -1), FD_P(0, 0, -1), GLT,
              FD_P(-1, 1, 0), FD_P(0, 1, 0), FD_P(-1, 0, 0), FD_P(-1, 1, -1), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(-1, 0, -1)),
        FD_T3(AXZ, -1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, -1), FD_P(0, 0, -1), GLT,
              FD_P(-1, 0, 0), FD_P(0, -1, 0), FD_P(-1, -1, 0), FD_P(-1, 0, -1), FD_P(0, 0, -1), FD_P(0, -1, -1),
              FD_P(-1, -1, -1)),
        FD_TNE(AXZ, 1, AXY, FD_P(0, 0, 1), GLT, FD_P(-1, -1, 1), FD_P(-1, 0, 1), FD_P(-1, 0, 0), FD_P(-1, 1, 1),
               FD_P(0, 0, 1), FD_DQ(1), FD_DH, FD_DS(1, 0), FD_DS(0, 0), FD_DH),
        FD_TNE(AXZ, -1, AXY, FD_P(0, 0, -1), GLT, FD_P(-1, 1, -1), FD_P(-1, 0, -1), FD_P(-1, 0, 0), FD_P(-1, -1, -1),
               FD_P(0, 0, -1), FD_DQ(-1), FD_DH, FD_DS(-1, 0), FD_DS(0, 0), FD_DH),
        FD_TNE(AXY, 1, AXZ, FD_P(0, 1, 0), GLT, FD_P(-1, 1, 1), FD_P(-1, 1, 0), FD_P(-1, 0, 0), FD_P(-1, 1, -1),
               FD_P(0, 1, 0), FD_DH, FD_DQ(0), FD_DS(0, 1), FD_DS(0, 0), FD_DQ(0)),
        FD_TNE(AXY, -1, AXZ, FD_P(0, -1, 0), GLT, FD_P(-1, -1, -1), FD_P(-1, -1, 0), FD_P(-1, 0, 0), FD_P(-1, -1, 1),
               FD_P(0, -1, 0), FD_DH, FD_DQ(0), FD_DS(0, -1), FD_DS(0, 0), FD_DQ(0)),
        FD_T2A(AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 0), FD_P(-1, 0, 0), FD_P(-1, -1, 0), FD_DQ(0), FD_DS(0, 0), FD_DS(0, -1),
               FD_DQ(0)),
        FD_T2A(AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 0), FD_P(-1, 0, 0), FD_P(-1, 1, 0), FD_DQ(0), FD_DS(0, 0), FD_DS(0, 1),
               FD_DQ(0)),
        FD_T2A(AXZ, -1, FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(-1, 0, 0), FD_P(-1, 0, -1), FD_DH, FD_DS(0, 0), FD_DS(-1, 0),
               FD_DH),
        FD_T2A(AXZ, 1, FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(-1, 0, 0), FD_P(-1, 0, 1), FD_DH, FD_DS(0, 0), FD_DS(1, 0), FD_DH),
        FD_T2B(AXZ, 1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 1), FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 0), FD_P(0, -1, 1),
               FD_DH, FD_DQ(0), FD_DQ(1), FD_DH),
        FD_T2B(AXZ, -1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, -1), FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(0, -1, 0),
               FD_P(0, -1, -1), FD_DH, FD_DQ(0), FD_DQ(-1), FD_DH),
        FD_T2B(AXZ, 1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(0, 1, 0), FD_P(0, 1, 1),
               FD_DH, FD_DQ(0), FD_DQ(1), FD_DH),
        FD_T2B(AXZ, -1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(0, 1, 0), FD_P(0, 1, -1),
               FD_DH, FD_DQ(0), FD_DQ(-1), FD_DH),
        FD_TNF(AXY, AXZ, FD_P(-1, 1, 0), FD_P(-1, 0, -1), FD_P(-1, 0, 0), FD_P(-1, 0, 1), FD_P(-1, -1, 0), FD_DQ(0), FD_DH,
               FD_DS(0, 0)),
        FD_T1P(FD_P(-1, 0, 0), FD_XR(0), FD_XS(0), FD_XF),
        FD_T1(AXY, 1, FD_P(0, 1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXY, -1, FD_P(0, -1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXZ, 1, FD_P(0, 0, 1), FD_XH, FD_D0, FD_D0),
        FD_T1(AXZ, -1, FD_P(0, 0, -1), FD_XH, FD_D0, FD_D0),
    };
};

// WEST: face x1, inner layer x1+1, sphfd.c:3275-4033.
template <>
struct FaceSpec<WEST> {
    static constexpr int nax = AXX, ndir = +1, a_ax = AXY, b_ax = AXZ;
    static constexpr Trial trials[22] = {
        FD_T3(AXZ, 1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 1), FD_P(0, 0, 1), GLT,
              FD_P(1, 1, 1), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(1, 0, 1), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0)),
        FD_T3(AXZ, 1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 1), FD_P(0, 0, 1), GLT,
              FD_P(1, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 1), FD_P(1, -1, 1), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0)),
        FD_T3(AXZ, -1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, -1), FD_P(0, 0, -1), GLT,
              FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0), FD_P(1, 1, -1), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(1, 0, -1)),
        FD_T3(AXZ, -1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, -1), FD_P(0, 0, -1), GLT,
              FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0), FD_P(1, 0, -1), FD_P(0, 0, -1), FD_P(0, -1, -1), FD_P(1, -1, -1)),
        FD_TNE(AXZ, 1, AXY, FD_P(0, 0, 1), GLT, FD_P(1, 1, 1), FD_P(1, 0, 1), FD_P(1, 0, 0), FD_P(1, -1, 1), FD_P(0, 0, 1),
               FD_DQ(1), FD_DH, FD_DS(1, 0), FD_DS(0, 0), FD_DH),
        FD_TNE(AXZ, -1, AXY, FD_P(0, 0, -1), GLT, FD_P(1, -1, -1), FD_P(1, 0, -1), FD_P(1, 0, 0), FD_P(1, 1, -1),
               FD_P(0, 0, -1), FD_DQ(-1), FD_DH, FD_DS(-1, 0), FD_DS(0, 0), FD_DH),
        FD_TNE(AXY, 1, AXZ, FD_P(0, 1, 0), GLT, FD_P(1, 1, -1), FD_P(1, 1, 0), FD_P(1, 0, 0), FD_P(1, 1, 1), FD_P(0, 1, 0),
               FD_DH, FD_DQ(0), FD_DS(0, 1), FD_DS(0, 0), FD_DQ(0)),
        FD_TNE(AXY, -1, AXZ, FD_P(0, -1, 0), GLT, FD_P(1, -1, 1), FD_P(1, -1, 0), FD_P(1, 0, 0), FD_P(1, -1, -1),
               FD_P(0, -1, 0), FD_DH, FD_DQ(0), FD_DS(0, -1), FD_DS(0, 0), FD_DQ(0)),
        FD_T2A(AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 0), FD_P(1, 0, 0), FD_P(1, -1, 0), FD_DQ(0), FD_DS(0, 0), FD_DS(0, -1),
               FD_DQ(0)),
        FD_T2A(AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0), FD_P(1, 1, 0), FD_DQ(0), FD_DS(0, 0), FD_DS(0, 1),
               FD_DQ(0)),
        FD_T2A(AXZ, -1, FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(1, 0, 0), FD_P(1, 0, -1), FD_DH, FD_DS(0, 0), FD_DS(-1, 0),
               FD_DH),
        FD_T2A(AXZ, 1, FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(1, 0, 0), FD_P(1, 0, 1), FD_DH, FD_DS(0, 0), FD_DS(1, 0), FD_DH),
        FD_T2B(AXZ, 1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, 1), FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 0), FD_P(0, -1, 1),
               FD_DH, FD_DQ(0), FD_DQ(1), FD_DH),
        FD_T2B(AXZ, -1, AXY, -1, FD_P(0, -1, 0), FD_P(0, -1, -1), FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(0, -1, 0),
               FD_P(0, -1, -1), FD_DH, FD_DQ(0), FD_DQ(-1), FD_DH),
        FD_T2B(AXZ, 1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(0, 1, 0), FD_P(0, 1, 1),
               FD_DH, FD_DQ(0), FD_DQ(1), FD_DH),
        FD_T2B(AXZ, -1, AXY, 1, FD_P(0, 1, 0), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(0, 1, 0), FD_P(0, 1, -1),
               FD_DH, FD_DQ(0), FD_DQ(-1), FD_DH),
        FD_TNF(AXY, AXZ, FD_P(1, 1, 0), FD_P(1, 0, 1), FD_P(1, 0, 0), FD_P(1, 0, -1), FD_P(1, -1, 0), FD_DQ(0), FD_DH,
               FD_DS(0, 0)),
        FD_T1P(FD_P(1, 0, 0), FD_XR(0), FD_XS(0), FD_XF),
        FD_T1(AXY, 1, FD_P(0, 1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXY, -1, FD_P(0, -1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXZ, 1, FD_P(0, 0, 1), FD_XH, FD_D0, FD_D0),
        FD_T1(AXZ, -1, FD_P(0, 0, -1), FD_XH, FD_D0, FD_D0),
    };
};

// NORTH: face y1, inner layer y1+1, in-face axes i (a) and k (b), sphfd.c:4035-4798.
template <>
struct FaceSpec<NORTH> {
    static constexpr int nax = AXY, ndir = +1, a_ax = AXX, b_ax = AXZ;
    static constexpr Trial trials[22] = {
        FD_T3(AXZ, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 1), FD_P(0, 0, 1), GLT,
              FD_P(1, 1, 1), FD_P(0, 1, 1), FD_P(0, 0, 1), FD_P(1, 0, 1), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0)),
        FD_T3(AXZ, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 1), FD_P(0, 0, 1), GLT,
              FD_P(0, 1, 1), FD_P(-1, 1, 1), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(0, 1, 0), FD_P(-1, 1, 0), FD_P(-1, 0, 0)),
        FD_T3(AXZ, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, -1), FD_P(0, 0, -1), GLT,
              FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 0, 0), FD_P(1, 1, -1), FD_P(0, 1, -1), FD_P(0, 0, -1), FD_P(1, 0, -1)),
        FD_T3(AXZ, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, -1), FD_P(0, 0, -1), GLT,
              FD_P(0, 1, 0), FD_P(-1, 1, 0), FD_P(-1, 0, 0), FD_P(0, 1, -1), FD_P(-1, 1, -1), FD_P(-1, 0, -1), FD_P(0, 0, -1)),
        FD_TNE(AXZ, 1, AXX, FD_P(0, 0, 1), GLT, FD_P(-1, 1, 1), FD_P(0, 1, 1), FD_P(0, 1, 0), FD_P(1, 1, 1), FD_P(0, 0, 1),
               FD_DS(1, 1), FD_DH, FD_DQ(1), FD_DQ(0), FD_DH),
        FD_TNE(AXZ, -1, AXX, FD_P(0, 0, -1), GLT, FD_P(1, 1, -1), FD_P(0, 1, -1), FD_P(0, 1, 0), FD_P(-1, 1, -1),
               FD_P(0, 0, -1), FD_DS(-1, 1), FD_DH, FD_DQ(-1), FD_DQ(0), FD_DH),
        FD_TNE(AXX, -1, AXZ, FD_P(-1, 0, 0), GLT, FD_P(-1, 1, -1), FD_P(-1, 1, 0), FD_P(0, 1, 0), FD_P(-1, 1, 1),
               FD_P(-1, 0, 0), FD_DH, FD_DS(0, 1), FD_DQ(0), FD_DQ(0), FD_DS(0, 0)),
        FD_TNE(AXX, 1, AXZ, FD_P(1, 0, 0), GLT, FD_P(1, 1, 1), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(1, 1, -1), FD_P(1, 0, 0),
               FD_DH, FD_DS(0, 1), FD_DQ(0), FD_DQ(0), FD_DS(0, 0)),
        FD_T2A(AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 0), FD_P(0, 1, 0), FD_P(1, 1, 0), FD_DS(0, 0), FD_DQ(0), FD_DQ(0), FD_DS(0, 1)),
        FD_T2A(AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 0), FD_P(0, 1, 0), FD_P(-1, 1, 0), FD_DS(0, 0), FD_DQ(0), FD_DQ(0),
               FD_DS(0, 1)),
        FD_T2A(AXZ, -1, FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(0, 1, 0), FD_P(0, 1, -1), FD_DH, FD_DQ(0), FD_DQ(-1), FD_DH),
        FD_T2A(AXZ, 1, FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(0, 1, 0), FD_P(0, 1, 1), FD_DH, FD_DQ(0), FD_DQ(1), FD_DH),
        FD_T2B(AXZ, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 1), FD_P(0, 0, 1), FD_P(1, 0, 0), FD_P(0, 0, 1), FD_P(1, 0, 1),
               FD_DS(0, 0), FD_DH, FD_DH, FD_DS(1, 0)),
        FD_T2B(AXZ, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, -1), FD_P(0, 0, -1), FD_P(1, 0, 0), FD_P(0, 0, -1), FD_P(1, 0, -1),
               FD_DS(0, 0), FD_DH, FD_DH, FD_DS(-1, 0)),
        FD_T2B(AXZ, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(-1, 0, 0), FD_P(0, 0, 1), FD_P(-1, 0, 1),
               FD_DS(0, 0), FD_DH, FD_DH, FD_DS(1, 0)),
        FD_T2B(AXZ, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, -1), FD_P(0, 0, -1), FD_P(-1, 0, 0), FD_P(0, 0, -1),
               FD_P(-1, 0, -1), FD_DS(0, 0), FD_DH, FD_DH, FD_DS(-1, 0)),
        FD_TNF(AXX, AXZ, FD_P(0, 1, 1), FD_P(1, 1, 0), FD_P(0, 1, 0), FD_P(-1, 1, 0), FD_P(0, 1, -1), FD_DH, FD_DS(0, 1),
               FD_DQ(0)),
        FD_T1P(FD_P(0, 1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXX, 1, FD_P(1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXX, -1, FD_P(-1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXZ, 1, FD_P(0, 0, 1), FD_XH, FD_D0, FD_D0),
        FD_T1(AXZ, -1, FD_P(0, 0, -1), FD_XH, FD_D0, FD_D0),
    };
};

// SOUTH: face y2, inner layer y2-1, sphfd.c:4800-5557.
template <>
struct FaceSpec<SOUTH> {
    static constexpr int nax = AXY, ndir = -1, a_ax = AXX, b_ax = AXZ;
    static constexpr Trial trials[22] = {
        FD_T3(AXZ, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 1), FD_P(0, 0, 1), GLT,
              FD_P(1, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 1), FD_P(1, -1, 1), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0)),
        FD_T3(AXZ, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 1), FD_P(0, 0, 1), GLT,
              FD_P(0, -1, 1), FD_P(-1, -1, 1), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 0), FD_P(-1, -1, 0), FD_P(-1, 0, 0)),
        FD_T3(AXZ, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, -1), FD_P(0, 0, -1), GLT,
              FD_P(1, -1, 0), FD_P(0, -1, 0), FD_P(1, 0, 0), FD_P(1, -1, -1), FD_P(0, -1, -1), FD_P(0, 0, -1), FD_P(1, 0, -1)),
        FD_T3(AXZ, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, -1), FD_P(0, 0, -1), GLT,
              FD_P(0, -1, 0), FD_P(-1, -1, 0), FD_P(-1, 0, 0), FD_P(0, -1, -1), FD_P(-1, -1, -1), FD_P(-1, 0, -1),
              FD_P(0, 0, -1)),
        FD_TNE(AXZ, 1, AXX, FD_P(0, 0, 1), GLT, FD_P(1, -1, 1), FD_P(0, -1, 1), FD_P(0, -1, 0), FD_P(-1, -1, 1),
               FD_P(0, 0, 1), FD_DS(1, -1), FD_DH, FD_DQ(1), FD_DQ(0), FD_DH),
        FD_TNE(AXZ, -1, AXX, FD_P(0, 0, -1), GLT, FD_P(-1, -1, -1), FD_P(0, -1, -1), FD_P(0, -1, 0), FD_P(1, -1, -1),
               FD_P(0, 0, -1), FD_DS(-1, -1), FD_DH, FD_DQ(-1), FD_DQ(0), FD_DH),
        FD_TNE(AXX, -1, AXZ, FD_P(-1, 0, 0), GLT, FD_P(-1, -1, 1), FD_P(-1, -1, 0), FD_P(0, -1, 0), FD_P(-1, -1, -1),
               FD_P(-1, 0, 0), FD_DH, FD_DS(0, -1), FD_DQ(0), FD_DQ(0), FD_DS(0, 0)),
        FD_TNE(AXX, 1, AXZ, FD_P(1, 0, 0), GLT, FD_P(1, -1, -1), FD_P(1, -1, 0), FD_P(0, -1, 0), FD_P(1, -1, 1),
               FD_P(1, 0, 0), FD_DH, FD_DS(0, -1), FD_DQ(0), FD_DQ(0), FD_DS(0, 0)),
        FD_T2A(AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 0), FD_P(0, -1, 0), FD_P(1, -1, 0), FD_DS(0, 0), FD_DQ(0), FD_DQ(0),
               FD_DS(0, -1)),
        FD_T2A(AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 0), FD_P(0, -1, 0), FD_P(-1, -1, 0), FD_DS(0, 0), FD_DQ(0), FD_DQ(0),
               FD_DS(0, -1)),
        FD_T2A(AXZ, -1, FD_P(0, 0, -1), FD_P(0, 0, -1), FD_P(0, -1, 0), FD_P(0, -1, -1), FD_DH, FD_DQ(0), FD_DQ(-1), FD_DH),
        FD_T2A(AXZ, 1, FD_P(0, 0, 1), FD_P(0, 0, 1), FD_P(0, -1, 0), FD_P(0, -1, 1), FD_DH, FD_DQ(0), FD_DQ(1), FD_DH),
        FD_T2B(AXZ, 1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, 1), FD_P(0, 0, 1), FD_P(1, 0, 0), FD_P(0, 0, 1), FD_P(1, 0, 1),
               FD_DS(0, 0), FD_DH, FD_DH, FD_DS(1, 0)),
        FD_T2B(AXZ, -1, AXX, 1, FD_P(1, 0, 0), FD_P(1, 0, -1), FD_P(0, 0, -1), FD_P(1, 0, 0), FD_P(0, 0, -1), FD_P(1, 0, -1),
               FD_DS(0, 0), FD_DH, FD_DH, FD_DS(-1, 0)),
        FD_T2B(AXZ, 1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, 1), FD_P(0, 0, 1), FD_P(-1, 0, 0), FD_P(0, 0, 1), FD_P(-1, 0, 1),
               FD_DS(0, 0), FD_DH, FD_DH, FD_DS(1, 0)),
        FD_T2B(AXZ, -1, AXX, -1, FD_P(-1, 0, 0), FD_P(-1, 0, -1), FD_P(0, 0, -1), FD_P(-1, 0, 0), FD_P(0, 0, -1),
               FD_P(-1, 0, -1), FD_DS(0, 0), FD_DH, FD_DH, FD_DS(-1, 0)),
        FD_TNF(AXX, AXZ, FD_P(0, -1, 1), FD_P(-1, -1, 0), FD_P(0, -1, 0), FD_P(1, -1, 0), FD_P(0, -1, -1), FD_DH,
               FD_DS(0, -1), FD_DQ(0)),
        FD_T1P(FD_P(0, -1, 0), FD_XR(0), FD_XQ, FD_D0),
        FD_T1(AXX, 1, FD_P(1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXX, -1, FD_P(-1, 0, 0), FD_XR(0), FD_XF, FD_XS(0)),
        FD_T1(AXZ, 1, FD_P(0, 0, 1), FD_XH, FD_D0, FD_D0),
        FD_T1(AXZ, -1, FD_P(0, 0, -1), FD_XH, FD_D0, FD_D0),
    };
};

// --------------------------------------------------------- geometry tables

// fdsph3d (sphfd.c:5944-5978) at node (j, k) of a cell corner: point 7 is the
// node, points 0..6 sit at offsets (dx, dy, dz) in the cell, and the signs
// g7*gi, n7*ni, m7*mi are +1 exactly when the point shares the node's layer,
// row or column (gi*gj, ni*nj, mi*mj likewise for two points).  Every
// coefficient of the quadratic is then
//     +-1/h^2 + (+-1/dq^2 + X) / R
// with X a reciprocal that depends on rows only and R = r_a * r_b a product
// of two layer radii.  X, R and RN(1/R) come from per-table row and layer
// tables, the division by R from xdiv, so each coefficient costs an add and
// one fast division and equals the reference's bit for bit.  The few values
// that depend on both row and layer (the reciprocal of the distance
// (r sin q) df and the `a` terms of the quadratics) are divided per node.
// Every entry is computed with the reference's operations in the reference's
// order.
//
// Row entry j (j = -1..ny), with sq(j) = sin(qc(j)), dfsq = df*df:
//   kRowIS2        1 / ((df sq(j)) (df sq(j)))               (D terms)
//   kRowZ + dy+1   1 / ((df (df sq(j))) sq(j+dy)), dy=-1..1  (b terms)
//   kRowX0         1 / ((dfsq sq(j)) sq(j))                  (pairs, same row)
//   kRowXp         1 / ((dfsq sq(j)) sq(j+1))                (pairs, rows j, j+1)
//   kRowXq         1 / ((dfsq sq(j+1)) sq(j))                (pairs, rows j+1, j)
// Layer entry k (k = -1..nz), r(k) = z0r - h k:
//   kLayR00, kLayY00   r(k) r(k) and its reciprocal
//   kLayR01, kLayY01   r(k) r(k+1) and its reciprocal
//   kLayYQ, kLayIQ2    RN(1 / Q) and RN(1 / (Q Q)), Q = r(k) dq
constexpr int kRowIS2 = 0, kRowZ = 1, kRowX0 = 4, kRowXp = 5, kRowXq = 6, kRowStride = 7;
constexpr int kLayR00 = 0, kLayY00 = 1, kLayR01 = 2, kLayY01 = 3, kLayYQ = 4, kLayIQ2 = 5, kLayStride = 6;

struct GeoParams {
    int ny, nz;
    double h, dq, df, z0r;
    const double *sinq;   // sin(qc(j)) for j = -1..ny, at [j + 1]
};

// Buffer layout: rows | layers.
FD_INLINE long long geo_rows_doubles(int ny) { return (long long)(ny + 2) * kRowStride; }
FD_INLINE long long geo_layers_doubles(int nz) { return (long long)(nz + 2) * kLayStride; }
FD_INLINE long long geo_doubles(int ny, int nz) { return geo_rows_doubles(ny) + geo_layers_doubles(nz); }
// Work items of a fill: one per row and layer entry.
FD_INLINE long long geo_items(int ny, int nz) { return (long long)(ny + 2) + (nz + 2); }

// Fills work item `it` of the geometry buffer `out`.
FD_INLINE void geo_fill_item(const GeoParams &G, long long it, double *out) {
    auto rcz = [&](int kk) { return G.z0r - G.h * (kk); };
    auto sq = [&](int jj) { return G.sinq[jj + 1]; };
    const double dfsq = G.df * G.df;
    const long long nr = G.ny + 2;
    if (it < nr) {
        const int j = (int)it - 1;
        double *E = out + it * kRowStride;
        for (int q = 0; q < kRowStride; q++) E[q] = 0.0;
        const double dfsinq = G.df * sq(j);
        E[kRowIS2] = 1. / (dfsinq * dfsinq);
        E[kRowX0] = 1. / (dfsq * sq(j) * sq(j));
        if (j + 1 <= G.ny) {
            E[kRowXp] = 1. / (dfsq * sq(j) * sq(j + 1));
            E[kRowXq] = 1. / (dfsq * sq(j + 1) * sq(j));
        }
        if (j >= 0 && j < G.ny) {
            const double df2sinq = G.df * dfsinq;
            for (int dy = -1; dy <= 1; dy++) E[kRowZ + dy + 1] = 1. / (df2sinq * sq(j + dy));
        }
        return;
    }
    it -= nr;
    const int k = (int)it - 1;
    double *E = out + nr * kRowStride + it * kLayStride;
    for (int q = 0; q < kLayStride; q++) E[q] = 0.0;
    E[kLayR00] = rcz(k) * rcz(k);
    E[kLayY00] = 1.0 / E[kLayR00];
    E[kLayR01] = rcz(k) * rcz(k + 1);
    E[kLayY01] = 1.0 / E[kLayR01];
    const double Q = rcz(k) * G.dq;
    E[kLayYQ] = 1.0 / Q;
    E[kLayIQ2] = 1.0 / (Q * Q);
}

// index of the distance pair {f, g} of the node (forms DH, DRDQ, DRSDF)
constexpr int geo_pair(int f, int g) {
    return ((f == DRSDF && g == DH) || (f == DH && g == DRSDF)) ? 0
           : ((f == DRDQ && g == DH) || (f == DH && g == DRDQ)) ? 1
                                                                : 2;
}

// --------------------------------------------------------- node evaluation

// Per-table geometry shared by all nodes (sphfd.c:151-156, 437-455).
// Pointer to the geometry row/layer tables, which live in LDS on the GPU.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(3))) double *GeoPtr;
#else
typedef const double *GeoPtr;
#endif

struct Geo {
    double h, dq, df, y0, z0r;
    double rh, ihsq, idqsq;   // RN(1/h), 1/(h h), 1/(dq dq)
    const double *sinq;       // sin(qc(j)) for j = -1..ny, stored at [j + 1]
    GeoPtr rowt;              // geometry row table in LDS (entry j at [(j + 1) * kRowStride])
    GeoPtr layt;              // geometry layer table in LDS (entry k at [(k + 1) * kLayStride])
    int nx, ny, nz;
    int box[3][2];            // current box sides x1 x2 / y1 y2 / z1 z2
};

FD_INLINE GeoPtr row_at(const Geo &G, int j) { return G.rowt + (j + 1) * kRowStride; }
FD_INLINE GeoPtr lay_at(const Geo &G, int k) { return G.layt + (k + 1) * kLayStride; }
// r(k+a) r(k+b) and its reciprocal (|a - b| <= 1); the product is commutative.
template <int A, int B>
FD_INLINE void layer_pair(const Geo &G, int k, double &R, double &Y) {
    if constexpr (A == B) {
        const GeoPtr L = lay_at(G, k + A);
        R = L[kLayR00];
        Y = L[kLayY00];
    } else {
        const GeoPtr L = lay_at(G, k + (A < B ? A : B));
        R = L[kLayR01];
        Y = L[kLayY01];
    }
}
// 1 / ((dfsq sq(j+a)) sq(j+b)) (|a - b| <= 1)
template <int A, int B>
FD_INLINE double row_pair(const Geo &G, int j) {
    if constexpr (A == B) return row_at(G, j + A)[kRowX0];
    else if constexpr (B == A + 1) return row_at(G, j + A)[kRowXp];
    else return row_at(G, j + B)[kRowXq];
}

FD_INLINE double rc(const Geo &G, int k) { return G.z0r - G.h * (k); }

// The 3x3x2 neighbourhood of a node: times and slownesses of the face plane
// (plane 0) and of the inner layer (plane 1), indexed plane*9 + (db+1)*3 + (da+1)
// along the face's in-plane axes (a, b).  For in-face neighbours the caller
// stores the final value of a neighbour that precedes the node in the face
// order and the pre-face value of one that follows it.  Entries outside the
// grid hold arbitrary values; no trial whose bounds pass reads them.
struct Nbhd {
    float t[18];
    float s[18];
    double sq[3];   // sin(qc(j-1)), sin(qc(j)), sin(qc(j+1))
};

template <int F>
constexpr int nb_index(int dx, int dy, int dz) {
    using S = FaceSpec<F>;
    const int o[3] = {dx, dy, dz};
    const int dn = o[S::nax], da = o[S::a_ax], db = o[S::b_ax];
    return (dn == 0 ? 0 : 9) + (db + 1) * 3 + (da + 1);
}

// Result of one node's trial sequence.
struct NodeOut {
    float tnew;          // value to store when `stored`
    float tout0, tout4;  // t[0], t[4] as the node leaves them (when `sets`)
    uint8_t stored;      // guess < old value
    uint8_t head;        // stored and fhead > 1e-3
    uint8_t selfset;     // t[0]/t[4] written by this node before its 2D trials
    uint8_t sets;        // t[0]/t[4] written by this node at all
    uint8_t used_in;     // a 2D trial ran on the carried t[0]/t[4]
};

template <int F, int TI>
FD_INLINE bool bounds_ok(const Geo &G, int i, int j, int k) {
    constexpr Trial T = FaceSpec<F>::trials[TI];
    const int c[3] = {i, j, k};
    const int n3[3] = {G.nx, G.ny, G.nz};
    if constexpr (T.bax0 >= 0) {
        if (!(T.bdir0 > 0 ? (c[T.bax0] < n3[T.bax0] - 1) : (c[T.bax0] > 0))) return false;
    }
    if constexpr (T.bax1 >= 0) {
        if (!(T.bdir1 > 0 ? (c[T.bax1] < n3[T.bax1] - 1) : (c[T.bax1] > 0))) return false;
    }
    if constexpr (T.iax0 >= 0) {
        if (!(c[T.iax0] > G.box[T.iax0][0] + 1 && c[T.iax0] < G.box[T.iax0][1] - 1)) return false;
    }
    if constexpr (T.iax1 >= 0) {
        if (!(c[T.iax1] > G.box[T.iax1][0] + 1 && c[T.iax1] < G.box[T.iax1][1] - 1)) return false;
    }
    return true;
}

template <int F, int TI>
FD_INLINE bool gates_ok(const Nbhd &N) {
    constexpr Trial T = FaceSpec<F>::trials[TI];
    bool ok = true;
    static_for<0, T.ngate>([&](auto G) {
        constexpr P3 g = T.gate[G];
        const float v = N.t[nb_index<F>(g.dx, g.dy, g.dz)];
        ok = ok && ((T.gtest[G] == GNZ) ? (v != 0.0f) : ((double)v < kBig));
    });
    return ok;
}

template <int F, int TI>
FD_INLINE bool trial_open(const Geo &G, const Nbhd &N, int i, int j, int k) {
    return bounds_ok<F, TI>(G, i, j, k) && gates_ok<F, TI>(N);
}

FD_INLINE double sinq_of(const Nbhd &N, int dy) { return N.sq[dy + 1]; }

FD_INLINE double dist_of(const Geo &G, const Nbhd &N, DS d, int k) {
    return d.form == DH ? G.h : d.form == DRDQ ? rc(G, k + d.dz) * G.dq : rc(G, k + d.dz) * sinq_of(N, d.dy) * G.df;
}
// A distance and its correctly rounded reciprocal: h and r dq from the
// tables, (r sin q) df divided here (its reciprocal depends on row and layer).
FD_INLINE Dist dist_y(const Geo &G, const Nbhd &N, DS d, int k) {
    Dist r;
    r.d = dist_of(G, N, d, k);
    r.y = d.form == DH ? G.rh : d.form == DRDQ ? lay_at(G, k + d.dz)[kLayYQ] : 1.0 / r.d;
    return r;
}

// a = 1/d^2 + 1/e^2 of fdsph2d/fdsphne for the node's distance pair P
// ({S,H}, {Q,H}, {S,Q}: S = (r sin q) df, Q = r dq, H = h) and RN(1/(2 a)).
template <int P>
FD_INLINE void pair_a(const Geo &G, const Nbhd &N, int k, double &a, double &ya) {
    const double iH = G.ihsq, iQ = lay_at(G, k)[kLayIQ2];
    if constexpr (P == 1) {
        a = iQ + iH;
    } else {
        const double S = dist_of(G, N, DS{DRSDF, 0, 0}, k);
        const double iS = 1.0 / (S * S);
        a = P == 0 ? iS + iH : iS + iQ;
    }
    ya = 1.0 / (2 * a);
}
FD_INLINE double factor_of(const Geo &G, const Nbhd &N, DS f, int k) {
    return f.form == XH ? G.h : f.form == XR ? rc(G, k + f.dz) : f.form == XDQ ? G.dq : f.form == XDF ? G.df
                                                                                             : sinq_of(N, f.dy);
}

// fdsph3d of trial TI of face F at node (j, k) (sphfd.c:5935-6010), every
// geometric sub-expression from the geometry tables.
template <int F, int TI>
FD_INLINE double st_3d(const double t[7], const double s[8], const Geo &G, int j, int k) {
    constexpr Trial T = FaceSpec<F>::trials[TI];
    double slo = 0.;
#pragma unroll
    for (int i = 0; i < 8; i++) slo += s[i];
    slo = .125 * slo;
    // b: (g7 gi)/h^2 + ((n7 ni)/dq^2 + (m7 mi)/(df (df sq7) sqi)) / (r7 ri)
    double b = 0;
    static_for<0, 7>([&](auto I) {
        constexpr P3 o = T.pt[I];
        double R, Y;
        layer_pair<0, o.dz>(G, k, R, Y);
        const double z = row_at(G, j)[kRowZ + o.dy + 1];
        const double w = (o.dy == 0 ? G.idqsq : -G.idqsq) + (o.dx == 0 ? z : -z);
        b += t[I] * ((o.dz == 0 ? G.ihsq : -G.ihsq) + xdiv_geo(w, R, Y));
    });
    b = 2.0 * b;
    // c: t_i^2 (1/h^2 + (1/dq^2 + 1/(df sqi)^2) / ri^2) ...
    double c = 0.;
    static_for<0, 7>([&](auto I) {
        constexpr P3 o = T.pt[I];
        const GeoPtr L = lay_at(G, k + o.dz);
        const double w = G.idqsq + row_at(G, j + o.dy)[kRowIS2];
        c += t[I] * t[I] * (G.ihsq + xdiv_geo(w, L[kLayR00], L[kLayY00]));
    });
    // ... + 2 t_i sum_j t_j ((gi gj)/h^2 + ((ni nj)/dq^2 + (mi mj)/((df^2 sqi) sqj)) / (ri rj))
    static_for<0, 6>([&](auto I) {
        constexpr P3 u = T.pt[I];
        double c2 = 0.;
        static_for<decltype(I)::value + 1, 7>([&](auto J) {
            constexpr P3 v = T.pt[J];
            double R, Y;
            layer_pair<u.dz, v.dz>(G, k, R, Y);
            const double x = row_pair<u.dy, v.dy>(G, j);
            const double w = (u.dy == v.dy ? G.idqsq : -G.idqsq) + (u.dx == v.dx ? x : -x);
            c2 += t[J] * ((u.dz == v.dz ? G.ihsq : -G.ihsq) + xdiv_geo(w, R, Y));
        });
        c += 2.0 * c2 * t[I];
    });
    // a: the node's own diagonal term
    const GeoPtr L7 = lay_at(G, k);
    const double a = G.ihsq + xdiv_geo(G.idqsq + row_at(G, j)[kRowIS2], L7[kLayR00], L7[kLayY00]);
    c = c - 16.0 * slo * slo;
    const double rad = b * b - 4.0 * a * c;
    if (!(rad >= 0.)) return kAcausal;
    const double sr = sqrt(rad);
    const double a2 = 2.0 * a, ya = 1.0 / (2.0 * a);
    const double x1 = xdiv(-b + sr, a2, ya);
    const double x2 = xdiv(-b - sr, a2, ya);
    bool bad1 = false, bad2 = false;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        bad1 |= (x1 < t[i]);
        bad2 |= (x2 < t[i]);
    }
    if (bad1) return bad2 ? kAcausal : x2;
    if (bad2) return x1;
    return (x1 < x2) ? x1 : x2;
}

// Whether the node will run a 2D trial on carried t[0]/t[4]: it depends only
// on gates and on the node's own pre-face value (the new-edge block runs only
// if guess > 1e9, and with no 3D trial guess is still the pre-face value).
template <int F>
FD_INLINE bool needs_carry(const Nbhd &N, const Geo &G, int i, int j, int k, float told) {
    bool any3 = false, anyne = false, any2 = false;
    static_for<0, 22>([&](auto TI) {
        constexpr Trial T = FaceSpec<F>::trials[TI];
        if constexpr (T.kind == K3D || T.kind == KNE || T.kind == K2D) {
            if (trial_open<F, TI>(G, N, i, j, k)) {
                if constexpr (T.kind == K3D) any3 = true;
                else if constexpr (T.kind == KNE) anyne = true;
                else any2 = true;
            }
        }
    });
    const bool self = any3 || ((double)told > kBig && anyne);
    return any2 && !self;
}

// One trial's stencil value (trial TI of face F at node (j, k)); for a 2D
// trial t0/t4 are the carried registers its causality test reads.  The
// arithmetic of every kind, in the reference's order; eval_node and the
// trial-parallel LEVEL rounds both evaluate trials through this.
template <int F, int TI>
FD_INLINE double trial_value(const Nbhd &N, const Geo &G, int j, int k, double t0, double t4) {
    constexpr Trial T = FaceSpec<F>::trials[TI];
    const float slow = N.s[nb_index<F>(0, 0, 0)];
    if constexpr (T.kind == K3D) {
        double t[7], s[8];
        static_for<0, 8>([&](auto P) {
            constexpr P3 o = P < 7 ? T.pt[P] : P3{0, 0, 0};
            constexpr int q = nb_index<F>(o.dx, o.dy, o.dz);
            if constexpr (P < 7) t[P] = N.t[q];
            s[P] = N.s[q];
        });
        return st_3d<F, TI>(t, s, G, j, k);
    } else if constexpr (T.kind == KNE) {
        double t[8], s[8];
        static_for<0, 5>([&](auto P) {
            constexpr int q = nb_index<F>(T.pt[P].dx, T.pt[P].dy, T.pt[P].dz);
            t[P] = N.t[q];
            s[P] = N.s[q];
        });
        s[5] = slow;
        static_assert(T.d[3].dz == 0 && T.d[3].dy == 0 && T.d[4].dz == 0 && T.d[4].dy == 0, "node distances");
        constexpr int pa = geo_pair(T.d[3].form, T.d[4].form);
        double a, ya;
        pair_a<pa>(G, N, k, a, ya);
        return st_ne(t, s, dist_y(G, N, T.d[0], k), dist_y(G, N, T.d[1], k), dist_y(G, N, T.d[2], k),
                     dist_y(G, N, T.d[3], k), dist_y(G, N, T.d[4], k), a, ya);
    } else if constexpr (T.kind == K2D) {
        double t[8], s[8];
        t[0] = t0;
        t[4] = t4;
        s[0] = slow;
        static_for<0, 3>([&](auto P) {
            constexpr int q = nb_index<F>(T.pt[P].dx, T.pt[P].dy, T.pt[P].dz);
            t[P + 1] = N.t[q];
            s[P + 1] = N.s[q];
        });
        static_assert(T.d[0].dz == 0 && T.d[0].dy == 0 && T.d[1].dz == 0 && T.d[1].dy == 0, "node distances");
        constexpr int pa = geo_pair(T.d[0].form, T.d[1].form);
        double a, ya;
        pair_a<pa>(G, N, k, a, ya);
        return st_2d(t, s, dist_y(G, N, T.d[0], k), dist_y(G, N, T.d[1], k), dist_y(G, N, T.d[2], k),
                     dist_y(G, N, T.d[3], k), a, ya);
    } else if constexpr (T.kind == KNF) {
        double t[8], s[8];
        static_for<0, 5>([&](auto P) {
            constexpr int q = nb_index<F>(T.pt[P].dx, T.pt[P].dy, T.pt[P].dz);
            t[P] = N.t[q];
            s[P] = N.s[q];
        });
        s[5] = slow;
        return st_nf(t, s, dist_y(G, N, T.d[0], k), dist_y(G, N, T.d[1], k), dist_of(G, N, T.d[2], k));
    } else {  // K1D
        constexpr P3 o = T.pt[0];
        constexpr int q = nb_index<F>(o.dx, o.dy, o.dz);
        double step = .5 * (slow + N.s[q]);   // float sum, then double
        static_for<0, 3>([&](auto Fi) {
            if constexpr (T.fac[Fi].form != XNONE) step = step * factor_of(G, N, T.fac[Fi], k);
        });
        return (double)N.t[q] + step;
    }
}

// The distance an improving head-wave trial divides by (fhead = (guess - tr)
// / (den * slow)): d01 of a 2D trial, the product of the step factors of a 1D
// trial.
template <int F, int TI>
FD_INLINE double trial_den(const Nbhd &N, const Geo &G, int k) {
    constexpr Trial T = FaceSpec<F>::trials[TI];
    if constexpr (T.kind == K2D) {
        return dist_of(G, N, T.d[0], k);
    } else {
        double den = 1.0;
        static_for<0, 3>([&](auto Fi) {
            if constexpr (T.fac[Fi].form != XNONE) {
                const double fv = factor_of(G, N, T.fac[Fi], k);
                den = (Fi == 0) ? fv : den * fv;
            }
        });
        return den;
    }
}

// t[0] and t[4] as a 3D, new-edge or new-face trial leaves them.
template <int F, int TI>
FD_INLINE void trial_t04(const Nbhd &N, float &t0, float &t4) {
    constexpr Trial T = FaceSpec<F>::trials[TI];
    t0 = N.t[nb_index<F>(T.pt[0].dx, T.pt[0].dy, T.pt[0].dz)];
    t4 = N.t[nb_index<F>(T.pt[4].dx, T.pt[4].dy, T.pt[4].dz)];
}

// The node's full trial sequence (e.g. TOP sphfd.c:1035-1760), on registers.
// tin0/tin4 are the t[0]/t[4] left by earlier nodes; used only if out.used_in.
template <int F>
FD_INLINE void eval_node(const Nbhd &N, const Geo &G, int i, int j, int k, float told, float tin0, float tin4,
                     NodeOut &out) {
    const float slow = N.s[nb_index<F>(0, 0, 0)];
    double t0 = tin0, t4 = tin4;   // the carried registers
    bool own04 = false, self = false, used_in = false;
    double guess = told;
    double fhead = 0.;
    bool ne_open = false;
    static_for<0, 22>([&](auto TI) {
        constexpr Trial T = FaceSpec<F>::trials[TI];
        if constexpr (TI == 4) ne_open = guess > kBig;   // the new-edge block guard is evaluated once
        if constexpr (T.kind == KNE) {
            if (!ne_open) return;
        }
        if constexpr (T.kind == KNF) {
            if (!(guess > kBig)) return;
        }
        if (!trial_open<F, TI>(G, N, i, j, k)) return;
        const double tr = trial_value<F, TI>(N, G, j, k, t0, t4);
        if constexpr (T.kind == K3D || T.kind == KNE || T.kind == KNF) {
            float a, b;
            trial_t04<F, TI>(N, a, b);
            t0 = a;
            t4 = b;
            own04 = true;
            if constexpr (T.kind != KNF) self = true;
        } else if constexpr (T.kind == K2D) {
            if (!own04) used_in = true;
        }
        if (tr < guess) {
            if constexpr (T.fhead) fhead = (guess - tr) / (trial_den<F, TI>(N, G, k) * (double)slow);
            guess = tr;
        }
    });
    out.stored = guess < (double)told;
    out.tnew = out.stored ? (float)guess : told;
    out.head = out.stored && fhead > kHeadTest;
    out.selfset = self;
    out.sets = own04;
    out.used_in = used_in;
    out.tout0 = (float)t0;
    out.tout4 = (float)t4;
}

}  // namespace fdsweep
