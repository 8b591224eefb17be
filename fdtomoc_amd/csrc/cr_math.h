// cr_math.h -- correctly rounded sine, cosine and two-argument arctangent
// in double precision for the angles of sphrayderv's ray positions (|x| up
// to a few radians), on the device and the host alike.
//
// A ray's next cell depends on comparisons of its position, so a last-bit
// difference in sin, cos or atan2 between the device library and the host
// C library the reference uses can send a ray down another cell.  The host
// library's results are the correctly rounded ones in all but rare cases;
// these functions produce the correctly rounded result by evaluating in
// double-double arithmetic (about 104 significant bits) and rounding once:
//
//  * sin/cos: x = k*pi/2 + r with pi/2 held as three doubles and the
//    products formed exactly with fused multiply-adds, |r| <= pi/4 in
//    double-double; Taylor series of sin(r)/r and cos(r) in r^2 by Horner's
//    rule with double-double coefficients 1/n! (terms to n = 29, below
//    2^-110 of the result for |r| <= pi/4; those below 2^-33 of it in
//    double, which keeps the error under 2^-85); the quadrant selects and
//    signs.
//  * atan2(y, x): a first approximation t0 from the library, then one
//    Newton step on y cos t - x sin t = 0 with sin t0, cos t0 from above:
//    t = t0 + (y cos t0 - x sin t0) / (x cos t0 + y sin t0), the numerator
//    formed exactly enough (the products' rounding errors by fused
//    multiply-adds), and rounded once.
//
// The constants are exact expansions of pi/2 and 1/n! (from rational
// arithmetic).  Must be compiled without FP contraction: the error-free
// transformations below rely on every add and multiply rounding separately.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CRM_INLINE __host__ __device__ __forceinline__
#else
#include <math.h>
#define CRM_INLINE inline __attribute__((always_inline))
#endif

namespace crm {

struct dd {
    double hi, lo;
};

CRM_INLINE dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

CRM_INLINE dd fast_two_sum(double a, double b) {   // |a| >= |b|
    const double s = a + b;
    return {s, b - (s - a)};
}

CRM_INLINE dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

CRM_INLINE dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}

CRM_INLINE dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}

// 1/n! as double-doubles, n = 0..29
#define CRM_F(h, l) dd{h, l}
CRM_INLINE dd inv_fact(int n) {
    switch (n) {
    case 0: return CRM_F(0x1p+0, 0.0);
    case 1: return CRM_F(0x1p+0, 0.0);
    case 2: return CRM_F(0x1p-1, 0.0);
    case 3: return CRM_F(0x1.5555555555555p-3, 0x1.5555555555555p-57);
    case 4: return CRM_F(0x1.5555555555555p-5, 0x1.5555555555555p-59);
    case 5: return CRM_F(0x1.1111111111111p-7, 0x1.1111111111111p-63);
    case 6: return CRM_F(0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65);
    case 7: return CRM_F(0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73);
    case 8: return CRM_F(0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76);
    case 9: return CRM_F(0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73);
    case 10: return CRM_F(0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76);
    case 11: return CRM_F(0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80);
    case 12: return CRM_F(0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83);
    case 13: return CRM_F(0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87);
    case 14: return CRM_F(0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92);
    case 15: return CRM_F(0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97);
    case 16: return CRM_F(0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101);
    case 17: return CRM_F(0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103);
    case 18: return CRM_F(0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107);
    case 19: return CRM_F(0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112);
    case 20: return CRM_F(0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120);
    case 21: return CRM_F(0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120);
    case 22: return CRM_F(0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124);
    case 23: return CRM_F(0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130);
    case 24: return CRM_F(0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135);
    case 25: return CRM_F(0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139);
    case 26: return CRM_F(0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143);
    case 27: return CRM_F(0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149);
    case 28: return CRM_F(0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153);
    default: return CRM_F(0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157);
    }
}
#undef CRM_F

// sin and cos of x (|x| <= 8) as double-doubles.
CRM_INLINE void dd_sincos(double x, dd &s, dd &c) {
    constexpr double P1 = 0x1.921fb54442d18p+0, P2 = 0x1.1a62633145c07p-54, P3 = -0x1.f1976b7ed8fbcp-110;
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    // r = x - k*(P1 + P2 + P3), the products exact
    const dd p1 = two_prod(k, P1);
    dd r = two_sum(x, -p1.hi);
    r.lo -= p1.lo;
    r = fast_two_sum(r.hi, r.lo);
    const dd p2 = two_prod(k, P2);
    r = dd_add(r, dd{-p2.hi, -p2.lo});
    r = dd_add(r, dd{-(k * P3), 0.0});
    const dd r2 = dd_mul(r, r);
    // sin(r)/r = sum (-1)^n r^2n / (2n+1)!, cos(r) = sum (-1)^n r^2n / (2n)!,
    // n = 0..14, by Horner's rule in r^2: the terms from n = 6 on (below
    // 2^-33 of the sum for |r| <= pi/4) in double, the others in
    // double-double
    const double s2 = r2.hi;
    double ts = inv_fact(29).hi, tc = inv_fact(28).hi;
    for (int n = 13; n >= 6; n--) {
        const double sg = (n & 1) ? -1.0 : 1.0;
        ts = ts * s2 + sg * inv_fact(2 * n + 1).hi;
        tc = tc * s2 + sg * inv_fact(2 * n).hi;
    }
    dd ps{ts, 0.0}, pc{tc, 0.0};
    for (int n = 5; n >= 0; n--) {
        const dd fs = inv_fact(2 * n + 1), fc = inv_fact(2 * n);
        const double sg = (n & 1) ? -1.0 : 1.0;
        ps = dd_add(dd_mul(ps, r2), dd{sg * fs.hi, sg * fs.lo});
        pc = dd_add(dd_mul(pc, r2), dd{sg * fc.hi, sg * fc.lo});
    }
    const dd sr = dd_mul(ps, r), cr = pc;
    const int q = ((int)k) & 3;
    switch (q) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = dd{-sr.hi, -sr.lo}; break;
    case 2: s = dd{-sr.hi, -sr.lo}; c = dd{-cr.hi, -cr.lo}; break;
    default: s = dd{-cr.hi, -cr.lo}; c = sr; break;
    }
}

CRM_INLINE double cr_sin(double x) {
    if (!(fabs(x) <= 8.0)) return sin(x);
    dd s, c;
    dd_sincos(x, s, c);
    return s.hi + s.lo;
}

CRM_INLINE double cr_cos(double x) {
    if (!(fabs(x) <= 8.0)) return cos(x);
    dd s, c;
    dd_sincos(x, s, c);
    return c.hi + c.lo;
}

CRM_INLINE void cr_sincos(double x, double &sv, double &cv) {
    if (!(fabs(x) <= 8.0)) {
        sv = sin(x);
        cv = cos(x);
        return;
    }
    dd s, c;
    dd_sincos(x, s, c);
    sv = s.hi + s.lo;
    cv = c.hi + c.lo;
}

// atan2(y, x); when the result equals the first approximation (most
// calls), also its sine and cosine, correctly rounded (have = true).
CRM_INLINE double cr_atan2_sc(double y, double x, double &sin_t, double &cos_t, bool &have) {
    have = false;
    const double t0 = atan2(y, x);
    if (!(fabs(t0) <= 8.0) || (x == 0.0 && y == 0.0) || !(fabs(x) < 1e300) || !(fabs(y) < 1e300)) return t0;
    dd s, c;
    dd_sincos(t0, s, c);
    // y cos t0 - x sin t0, the leading products' errors kept
    const dd a = two_prod(y, c.hi), b = two_prod(x, s.hi);
    const double num = (a.hi - b.hi) + ((a.lo - b.lo) + (y * c.lo - x * s.lo));
    const double den = x * c.hi + y * s.hi;
    if (!(den > 0.0)) return t0;
    const double t = t0 + num / den;
    if (t == t0) {
        sin_t = s.hi + s.lo;
        cos_t = c.hi + c.lo;
        have = true;
    }
    return t;
}

CRM_INLINE double cr_atan2(double y, double x) {
    double s, c;
    bool have;
    return cr_atan2_sc(y, x, s, c, have);
}

}  // namespace crm
