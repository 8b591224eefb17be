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
    for (int i = 0; i < 8; i++) slo += s[i];
    slo = .125 * slo;
    const double hsq = h * h, dqsq = dq * dq, dfsq = df * df;
    const double dfsinq7 = df * sq[7];
    const double df2sinq7 = df * dfsinq7;
    const double a = 1.0 / hsq + (1.0 / dqsq + 1. / (dfsinq7 * dfsinq7)) / (r[7] * r[7]);
    double b = 0;
    for (int i = 0; i < 7; i++)
        b += t[i] * ((double)(g[7] * g[i]) / hsq +
                     ((double)(n[7] * n[i]) / dqsq + (double)(m[7] * m[i]) / (df2sinq7 * sq[i])) / (r[7] * r[i]));
    b = 2.0 * b;
    double c = 0.;
    for (int i = 0; i < 7; i++) {
        const double dfsinqi = df * sq[i];
        c += t[i] * t[i] * (1. / hsq + (1. / dqsq + 1. / (dfsinqi * dfsinqi)) / (r[i] * r[i]));
    }
    for (int i = 0; i < 6; i++) {
        double c2 = 0.;
        const double df2sinqi = dfsq * sq[i];
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
FD_INLINE double qcol(const Geo &G, int j) { return G.y0 + G.dq * (j); }
FD_INLINE double sinq(const Geo &G, int j) { return G.sinq[j + 1]; }

FD_INLINE double dist_of(const Geo &G, DS d, int j, int k) {
    return d.form == DH ? G.h : d.form == DRDQ ? rc(G, k + d.dz) * G.dq : rc(G, k + d.dz) * sinq(G, j + d.dy) * G.df;
}
FD_INLINE double factor_of(const Geo &G, DS f, int j, int k) {
    return f.form == XH ? G.h : f.form == XR ? rc(G, k + f.dz) : f.form == XDQ ? G.dq : f.form == XDF ? G.df : sinq(G, j + f.dy);
}

// Result of one node's trial sequence.
struct NodeOut {
    float tnew;        // value to store when `stored`
    float tout0, tout4;  // t[0], t[4] as the node leaves them (when `sets`)
    uint8_t stored;    // guess < old value
    uint8_t head;      // stored and fhead > 1e-3
    uint8_t selfset;   // t[0]/t[4] written by this node before its 2D trials
    uint8_t sets;      // t[0]/t[4] written by this node at all
    uint8_t used_in;   // a 2D trial ran on the carried t[0]/t[4]
};

// Accessor contract:
//   float t(dx,dy,dz)  time at a neighbour; for in-face neighbours the caller
//                       returns the final value if it precedes the node in the
//                       face order and the pre-face value otherwise.
//   float s(dx,dy,dz)  slowness.
template <int F, class Acc>
FD_INLINE bool bounds_ok(const Trial &T, const Geo &G, const int c[3]) {
    const int n3[3] = {G.nx, G.ny, G.nz};
    if (T.bax0 >= 0 && !(T.bdir0 > 0 ? (c[T.bax0] < n3[T.bax0] - 1) : (c[T.bax0] > 0))) return false;
    if (T.bax1 >= 0 && !(T.bdir1 > 0 ? (c[T.bax1] < n3[T.bax1] - 1) : (c[T.bax1] > 0))) return false;
    if (T.iax0 >= 0 && !(c[T.iax0] > G.box[T.iax0][0] + 1 && c[T.iax0] < G.box[T.iax0][1] - 1)) return false;
    if (T.iax1 >= 0 && !(c[T.iax1] > G.box[T.iax1][0] + 1 && c[T.iax1] < G.box[T.iax1][1] - 1)) return false;
    return true;
}

template <int F, class Acc>
FD_INLINE bool gates_ok(const Trial &T, const Acc &A) {
    for (int g = 0; g < T.ngate; g++) {
        const float v = A.t(T.gate[g].dx, T.gate[g].dy, T.gate[g].dz);
        const bool ok = (T.gtest[g] == GNZ) ? (v != 0.0f) : ((double)v < kBig);
        if (!ok) return false;
    }
    return true;
}

// Whether the node will run a 2D trial on carried t[0]/t[4]: it depends only
// on gates and on the node's own pre-face value (the new-edge block runs only
// if guess > 1e9, and with no 3D trial guess is still the pre-face value).
template <int F, class Acc>
FD_HD bool needs_carry(const Acc &A, const Geo &G, int i, int j, int k, float told) {
    using S = FaceSpec<F>;
    const int c[3] = {i, j, k};
    bool any3 = false, anyne = false, any2 = false;
    for (int ti = 0; ti < 22; ti++) {
        const Trial &T = S::trials[ti];
        if (T.kind == KNF || T.kind == K1D) continue;
        if (!bounds_ok<F, Acc>(T, G, c) || !gates_ok<F, Acc>(T, A)) continue;
        if (T.kind == K3D) any3 = true;
        else if (T.kind == KNE) anyne = true;
        else any2 = true;
    }
    const bool self = any3 || ((double)told > kBig && anyne);
    return any2 && !self;
}

// The node's full trial sequence (e.g. TOP sphfd.c:1035-1760).  tin0/tin4 are
// the t[0]/t[4] left by earlier nodes; used only if out.used_in.
template <int F, class Acc>
FD_HD void eval_node(const Acc &A, const Geo &G, int i, int j, int k, float told, float slow, float tin0,
                     float tin4, NodeOut &out) {
    using S = FaceSpec<F>;
    const int c[3] = {i, j, k};
    double t[8];
    double s[8];
    t[0] = tin0;
    t[4] = tin4;
    bool own04 = false, self = false;
    double guess = told;
    double fhead = 0.;
    bool ne_open = false;
    out.used_in = 0;
#pragma unroll
    for (int ti = 0; ti < 22; ti++) {
        const Trial &T = S::trials[ti];
        if (ti == 4) ne_open = guess > kBig;   // the new-edge block guard is evaluated once
        if (T.kind == KNE && !ne_open) continue;
        if (T.kind == KNF && !(guess > kBig)) continue;
        if (!bounds_ok<F, Acc>(T, G, c) || !gates_ok<F, Acc>(T, A)) continue;
        double tr;
        if (T.kind == K3D) {
            double r[8], sq[8];
            int g[8], n[8], m[8];
            int zmax = 0, ymax = 0, xmax = 0;
            for (int p = 0; p < 7; p++) {
                zmax = T.pt[p].dz > zmax ? T.pt[p].dz : zmax;
                ymax = T.pt[p].dy > ymax ? T.pt[p].dy : ymax;
                xmax = T.pt[p].dx > xmax ? T.pt[p].dx : xmax;
            }
            for (int p = 0; p < 8; p++) {
                const P3 o = p < 7 ? T.pt[p] : P3{0, 0, 0};
                if (p < 7) {
                    t[p] = A.t(o.dx, o.dy, o.dz);
                    s[p] = A.s(o.dx, o.dy, o.dz);
                } else {
                    s[p] = slow;
                }
                r[p] = rc(G, k + o.dz);
                sq[p] = sinq(G, j + o.dy);
                // +1 on the cell's smaller-z layer, larger-y row, larger-x column
                g[p] = (o.dz == zmax) ? -1 : 1;
                n[p] = (o.dy == ymax) ? 1 : -1;
                m[p] = (o.dx == xmax) ? 1 : -1;
            }
            tr = st_3d(t, s, r, sq, g, n, m, G.h, G.dq, G.df);
            own04 = true;
            self = true;
            if (tr < guess) guess = tr;
        } else if (T.kind == KNE) {
            for (int p = 0; p < 5; p++) {
                t[p] = A.t(T.pt[p].dx, T.pt[p].dy, T.pt[p].dz);
                s[p] = A.s(T.pt[p].dx, T.pt[p].dy, T.pt[p].dz);
            }
            s[5] = slow;
            tr = st_ne(t, s, dist_of(G, T.d[0], j, k), dist_of(G, T.d[1], j, k), dist_of(G, T.d[2], j, k),
                       dist_of(G, T.d[3], j, k), dist_of(G, T.d[4], j, k));
            own04 = true;
            self = true;
            if (tr < guess) guess = tr;
        } else if (T.kind == K2D) {
            s[0] = slow;
            for (int p = 0; p < 3; p++) {
                t[p + 1] = A.t(T.pt[p].dx, T.pt[p].dy, T.pt[p].dz);
                s[p + 1] = A.s(T.pt[p].dx, T.pt[p].dy, T.pt[p].dz);
            }
            if (!own04) out.used_in = 1;
            const double d01 = dist_of(G, T.d[0], j, k);
            tr = st_2d(t, s, d01, dist_of(G, T.d[1], j, k), dist_of(G, T.d[2], j, k), dist_of(G, T.d[3], j, k));
            if (tr < guess) {
                if (T.fhead) fhead = (guess - tr) / (d01 * (double)slow);
                guess = tr;
            }
        } else if (T.kind == KNF) {
            for (int p = 0; p < 5; p++) {
                t[p] = A.t(T.pt[p].dx, T.pt[p].dy, T.pt[p].dz);
                s[p] = A.s(T.pt[p].dx, T.pt[p].dy, T.pt[p].dz);
            }
            s[5] = slow;
            tr = st_nf(t, s, dist_of(G, T.d[0], j, k), dist_of(G, T.d[1], j, k), dist_of(G, T.d[2], j, k));
            own04 = true;
            if (tr < guess) guess = tr;
        } else {  // K1D
            const P3 o = T.pt[0];
            double step = .5 * (slow + A.s(o.dx, o.dy, o.dz));   // float sum, then double
            double den = 1.0;
            for (int f = 0; f < 3; f++) {
                if (T.fac[f].form == XNONE) break;
                const double fv = factor_of(G, T.fac[f], j, k);
                step = step * fv;
                den = (f == 0) ? fv : den * fv;
            }
            tr = (double)A.t(o.dx, o.dy, o.dz) + step;
            if (tr < guess) {
                if (T.fhead) fhead = (guess - tr) / (den * (double)slow);
                guess = tr;
            }
        }
    }
    out.stored = guess < (double)told;
    out.tnew = out.stored ? (float)guess : told;
    out.head = out.stored && fhead > kHeadTest;
    out.selfset = self;
    out.sets = own04;
    out.tout0 = (float)t[0];
    out.tout4 = (float)t[4];
}

}  // namespace fdsweep
