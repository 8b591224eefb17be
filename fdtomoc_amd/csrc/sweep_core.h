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
#else
#include <math.h>
#define FD_HD
#define FD_INLINE inline __attribute__((always_inline))
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

// ---------------------------------------------------------------- stencils

// 8-point transmission stencil.  s holds 8 slownesses, r/q radii and
// colatitudes, sq[] = sin(q[]) (host libm values), g/n/m the +-1 signs.
FD_INLINE double st_3d(const double t[8], const double s[8], const double r[8], const double sq[8],
                       const int g[8], const int n[8], const int m[8], double h, double dq, double df) {
    double slo = 0.;
#pragma unroll
    for (int i = 0; i < 8; i++) slo += s[i];
    slo = .125 * slo;
    const double hsq = h * h, dqsq = dq * dq, dfsq = df * df;
    const double dfsinq7 = df * sq[7];
    const double df2sinq7 = df * dfsinq7;
    const double a = 1.0 / hsq + (1.0 / dqsq + 1. / (dfsinq7 * dfsinq7)) / (r[7] * r[7]);
    double b = 0;
#pragma unroll
    for (int i = 0; i < 7; i++)
        b += t[i] * ((double)(g[7] * g[i]) / hsq +
                     ((double)(n[7] * n[i]) / dqsq + (double)(m[7] * m[i]) / (df2sinq7 * sq[i])) / (r[7] * r[i]));
    b = 2.0 * b;
    double c = 0.;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const double dfsinqi = df * sq[i];
        c += t[i] * t[i] * (1. / hsq + (1. / dqsq + 1. / (dfsinqi * dfsinqi)) / (r[i] * r[i]));
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double c2 = 0.;
        const double df2sinqi = dfsq * sq[i];
#pragma unroll
        for (int j = i + 1; j < 7; j++)
            c2 += t[j] * ((double)(g[i] * g[j]) / hsq +
                          ((double)(n[i] * n[j]) / dqsq + (double)(m[i] * m[j]) / (df2sinqi * sq[j])) / (r[i] * r[j]));
        c += 2.0 * c2 * t[i];
    }
    c = c - 16.0 * slo * slo;
    const double rad = b * b - 4.0 * a * c;
    if (!(rad >= 0.)) return kAcausal;
    const double sr = sqrt(rad);
    const double x1 = (-b + sr) / (2.0 * a);
    const double x2 = (-b - sr) / (2.0 * a);
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

// New-edge stencil: points 0..4 known, point 5 the node.
FD_INLINE double st_ne(const double t[8], const double s[8], double d01, double d12, double d14, double d25,
                       double d45) {
    const double slo = .25 * (s[1] + s[2] + s[4] + s[5]);
    const double dt1 = (t[4] - t[1]) / d14 - t[2] / d25;
    const double dt2 = (t[2] - t[1]) / d12 - t[4] / d45;
    const double dt3 = (t[0] - t[3]) / d01;
    const double a = 1. / (d25 * d25) + 1. / (d45 * d45);
    const double b = 2 * (dt1 / d25 + dt2 / d45);
    const double c = dt1 * dt1 + dt2 * dt2 + dt3 * dt3 - 4 * slo * slo;
    const double rad = b * b - 4.0 * a * c;
    if (!(rad >= 0.)) return kAcausal;
    const double sr = sqrt(rad);
    const double x1 = (-b + sr) / (2 * a);
    const double x2 = (-b - sr) / (2 * a);
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

// 2D face stencil.  Causality is tested against t[0] and t[4] as well, which
// the caller carries over from earlier trials (the reference's 2D call sites
// never set them).
FD_INLINE double st_2d(const double t[8], const double s[8], double d01, double d02, double d13, double d23) {
    const double slo = .25 * (s[0] + s[1] + s[2] + s[3]);
    const double d01sq = d01 * d01, d02sq = d02 * d02;
    const double t1mt3 = t[1] - t[3], t2mt3 = t[2] - t[3];
    const double t1t3t2 = t1mt3 / d13 - t[2] / d02;
    const double t2t3t1 = t2mt3 / d23 - t[1] / d01;
    const double a = (1.0 / d01sq + 1. / d02sq);
    const double b = 2.0 * (t1t3t2 / d02 + t2t3t1 / d01);
    const double c = t1t3t2 * t1t3t2 + t2t3t1 * t2t3t1 - 4 * slo * slo;
    const double rad = b * b - 4.0 * a * c;
    if (!(rad >= 0.)) return kAcausal;
    const double sr = sqrt(rad);
    const double x1 = (-b + sr) / (2 * a);
    const double x2 = (-b - sr) / (2 * a);
    const bool bad1 = (x1 < t[1]) | (x1 < t[2]) | (x1 < t[3]) | (x1 < t[4]) | (x1 < t[0]);
    const bool bad2 = (x2 < t[1]) | (x2 < t[2]) | (x2 < t[3]) | (x2 < t[4]) | (x2 < t[0]);
    if (bad1) return bad2 ? kAcausal : x2;
    if (bad2) return x1;
    return (x2 < x1) ? x2 : x1;
}

// New-face stencil: explicit step from the inner layer.
FD_INLINE double st_nf(const double t[8], const double s[8], double d02, double d12, double d25) {
    const double slo = .5 * (s[5] + s[2]);
    const double t4t0 = (t[4] - t[0]) / d02;
    const double t1t3 = (t[1] - t[3]) / d12;
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

// --------------------------------------------------------- node evaluation

// Per-table geometry shared by all nodes (sphfd.c:151-156, 437-455).
struct Geo {
    double h, dq, df, y0, z0r;
    const double *sinq;   // sin(qc(j)) for j = -1..ny, stored at [j + 1]
    int nx, ny, nz;
    int box[3][2];        // current box sides x1 x2 / y1 y2 / z1 z2
};

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

template <int I, int N, class Fn>
FD_INLINE void static_for(Fn &&fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
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
FD_INLINE double factor_of(const Geo &G, const Nbhd &N, DS f, int k) {
    return f.form == XH ? G.h : f.form == XR ? rc(G, k + f.dz) : f.form == XDQ ? G.dq : f.form == XDF ? G.df
                                                                                             : sinq_of(N, f.dy);
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
        if constexpr (T.kind == K3D) {
            double t[8], s[8], r[8], sq[8];
            int g[8], n[8], m[8];
            constexpr int zmax = (T.pt[0].dz > 0 || T.pt[1].dz > 0 || T.pt[2].dz > 0 || T.pt[3].dz > 0 ||
                                  T.pt[4].dz > 0 || T.pt[5].dz > 0 || T.pt[6].dz > 0) ? 1 : 0;
            constexpr int ymax = (T.pt[0].dy > 0 || T.pt[1].dy > 0 || T.pt[2].dy > 0 || T.pt[3].dy > 0 ||
                                  T.pt[4].dy > 0 || T.pt[5].dy > 0 || T.pt[6].dy > 0) ? 1 : 0;
            constexpr int xmax = (T.pt[0].dx > 0 || T.pt[1].dx > 0 || T.pt[2].dx > 0 || T.pt[3].dx > 0 ||
                                  T.pt[4].dx > 0 || T.pt[5].dx > 0 || T.pt[6].dx > 0) ? 1 : 0;
            static_for<0, 8>([&](auto P) {
                constexpr P3 o = P < 7 ? T.pt[P] : P3{0, 0, 0};
                constexpr int q = nb_index<F>(o.dx, o.dy, o.dz);
                if constexpr (P < 7) t[P] = N.t[q];
                s[P] = N.s[q];
                r[P] = rc(G, k + o.dz);
                sq[P] = sinq_of(N, o.dy);
                // +1 on the cell's smaller-z layer, larger-y row, larger-x column
                g[P] = (o.dz == zmax) ? -1 : 1;
                n[P] = (o.dy == ymax) ? 1 : -1;
                m[P] = (o.dx == xmax) ? 1 : -1;
            });
            const double tr = st_3d(t, s, r, sq, g, n, m, G.h, G.dq, G.df);
            t0 = t[0];
            t4 = t[4];
            own04 = true;
            self = true;
            if (tr < guess) guess = tr;
        } else if constexpr (T.kind == KNE) {
            double t[8], s[8];
            static_for<0, 5>([&](auto P) {
                constexpr int q = nb_index<F>(T.pt[P].dx, T.pt[P].dy, T.pt[P].dz);
                t[P] = N.t[q];
                s[P] = N.s[q];
            });
            s[5] = slow;
            const double tr = st_ne(t, s, dist_of(G, N, T.d[0], k), dist_of(G, N, T.d[1], k), dist_of(G, N, T.d[2], k),
                                    dist_of(G, N, T.d[3], k), dist_of(G, N, T.d[4], k));
            t0 = t[0];
            t4 = t[4];
            own04 = true;
            self = true;
            if (tr < guess) guess = tr;
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
            if (!own04) used_in = true;
            const double d01 = dist_of(G, N, T.d[0], k);
            const double tr = st_2d(t, s, d01, dist_of(G, N, T.d[1], k), dist_of(G, N, T.d[2], k), dist_of(G, N, T.d[3], k));
            if (tr < guess) {
                if constexpr (T.fhead) fhead = (guess - tr) / (d01 * (double)slow);
                guess = tr;
            }
        } else if constexpr (T.kind == KNF) {
            double t[8], s[8];
            static_for<0, 5>([&](auto P) {
                constexpr int q = nb_index<F>(T.pt[P].dx, T.pt[P].dy, T.pt[P].dz);
                t[P] = N.t[q];
                s[P] = N.s[q];
            });
            s[5] = slow;
            const double tr = st_nf(t, s, dist_of(G, N, T.d[0], k), dist_of(G, N, T.d[1], k), dist_of(G, N, T.d[2], k));
            t0 = t[0];
            t4 = t[4];
            own04 = true;
            if (tr < guess) guess = tr;
        } else {  // K1D
            constexpr P3 o = T.pt[0];
            constexpr int q = nb_index<F>(o.dx, o.dy, o.dz);
            double step = .5 * (slow + N.s[q]);   // float sum, then double
            double den = 1.0;
            static_for<0, 3>([&](auto Fi) {
                if constexpr (T.fac[Fi].form != XNONE) {
                    const double fv = factor_of(G, N, T.fac[Fi], k);
                    step = step * fv;
                    den = (Fi == 0) ? fv : den * fv;
                }
            });
            const double tr = (double)N.t[q] + step;
            if (tr < guess) {
                if constexpr (T.fhead) fhead = (guess - tr) / (den * (double)slow);
                guess = tr;
            }
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
