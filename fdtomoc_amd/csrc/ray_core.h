// ray_core.h -- one ray of sphrayderv's ray tracing (src/sphrayderv.c: event
// interpolation 1306-1445, ray loop 1466-2250, lib/sphrayderv/dfind.c and
// pmin.c) as a function of one lane: the ray walks from the event back to
// the station's cell down the travel-time gradient of the station's table,
// cell by cell, and adds each segment's length to the coarse-grid
// derivatives at its midpoint through an accumulator.
//
// Shared by the gfx950 kernel (ray_kernel.hip) and a host build that checks
// its logic against the CPU restatement on the CPU.  Arithmetic follows the
// reference's declared types and order: positions, directions and step
// lengths in double; table differences, trilinear weights, midpoints and
// derivatives in float where the reference stores floats.  Every sine and
// cosine of a cell face -- a function of the cell's row or column only --
// comes from tables the host fills with its C library (the reference's own
// values); the lane evaluates only those of the ray's position, through a
// math policy: correctly rounded sine, cosine and arctangent on the device
// (cr_math.h; the host library's results in all but about one call in a
// thousand, where the host's are not correctly rounded), the host library
// itself in the host build.  Must be compiled without FP contraction.
#pragma once

#include <limits.h>
#include <stdint.h>

#include "cr_math.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RAY_HD __host__ __device__
#define RAY_INLINE __host__ __device__ __forceinline__
#else
#include <math.h>
#define RAY_HD
#define RAY_INLINE inline __attribute__((always_inline))
#endif

namespace fdray {

// sphrayderv.c:96-97: degrad is a float literal held in a double, rearth a float.
constexpr double kDegRad = (double)0.017453292f;
constexpr float kREarth = 6371.0f;

enum : int {
    kRayOk = 1,           // reached the station cell
    kRayUnreached = 0,    // a corner of the event's cell holds no time (> 1e9)
    kRayLeft = -1,        // left the medium
    kRayEventOut = -2,    // event outside the grid
    kRayLimit = -3,       // more than max_segments segments
    kRayCoarse = -4,      // a segment midpoint beyond the coarse grid (the reference asserts)
    kRayEdge = -5         // a cell whose corners lie outside the table
};

// The grid in sphrayderv's units (radians, geocentric colatitude, km).
// trig: per row j in [-1, ny] four doubles sin(yj), cos(yj), sin(yj + dq),
// cos(yj + dq) with yj = dq*j + y0, then per column i in [-1, nx]
// sin(xi), cos(xi), sin(xi + df), cos(xi + df) with xi = df*i + x0.
// mx/my/mz: per fine node index, a first guess of the coarse search's
// result near it (any guess gives the same result; a close one saves steps).
struct RayGrid {
    int nx, ny, nz, nxc, nyc, nzc;
    int ido1d, max_segments;
    double x0, y0, z0, df, dq, h;
    const float *gx, *gy, *gz;
    const double *trig;
    const int *mx, *my, *mz;
};

// One ray: event and station (sphrayderv.c:949-951, 874-876), and the
// station's Cartesian point as floats (the straight-ray target inside the
// station cube, sphrayderv.c:1833-1836, a function of the station only).
struct RayIn {
    double ex, ey, ez, xs, ys, zs;
    float x2, y2, z2;
    int table;   // index into the launch's table pointers
};

// sf0/sq0/sr0: the first step's unit gradient direction (longitude,
// colatitude, radius), from which the host computes the take-off angles with
// its own atan2f (sphrayderv.c:1865-1873; the device's atan2f rounds
// differently); not set when first_in_cube.
struct RayOut {
    int status, nseg, ndu, first_in_cube;
    double tev;
    float dtdx, dtdy, dtdz, rdev;
    double sf0, sq0, sr0;
};

// The position's sine, cosine and arctangent: correctly rounded.
// atan2_sc also hands back the result's sine and cosine when it has them.
struct CrMath {
    static RAY_INLINE void sincos(double x, double &s, double &c) { crm::cr_sincos(x, s, c); }
    static RAY_INLINE double atan2_sc(double y, double x, double &s, double &c, bool &have) {
        return crm::cr_atan2_sc(y, x, s, c, have);
    }
};

// ... or those of the C library in use (the host build of the tracer).
struct LibMath {
    static RAY_INLINE void sincos(double x, double &s, double &c) {
        s = ::sin(x);
        c = ::cos(x);
    }
    static RAY_INLINE double atan2_sc(double y, double x, double &, double &, bool &have) {
        have = false;
        return ::atan2(y, x);
    }
};

RAY_INLINE int ray_to_int(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
    return (int)v;
}

RAY_INLINE double ray_pmin(double d1, double d2) {
    if (d1 >= 0) return (d2 >= 0 && d2 <= d1) ? d2 : d1;
    return d2 >= 0 ? d2 : -1e20;
}

// dfind: the distances ordered by magnitude (ties to the lower axis), the
// smallest unless below tolmin, then the middle, then the largest.
RAY_INLINE void ray_pick(const double dd[3], double tolmin, double tolmax, double &d, int &md) {
    int mn, mid, mx;
    const double a0 = fabs(dd[0]), a1 = fabs(dd[1]), a2 = fabs(dd[2]);
    if (a0 <= a1 && a0 <= a2) {
        mn = 0;
        if (a1 <= a2) { mid = 1; mx = 2; } else { mid = 2; mx = 1; }
    } else if (a1 <= a2) {
        mn = 1;
        if (a0 <= a2) { mid = 0; mx = 2; } else { mid = 2; mx = 0; }
    } else {
        mn = 2;
        if (a0 <= a1) { mid = 0; mx = 1; } else { mid = 1; mx = 0; }
    }
    d = dd[mn];
    md = mn;
    if (dd[mn] < tolmin) {
        if (dd[mid] < tolmin) {
            if (dd[mx] < tolmax) { d = dd[mx]; md = mx; }
            else { d = dd[mid]; md = mid; }
        } else if (dd[mid] < tolmax) {
            d = dd[mid];
            md = mid;
        }
    }
}

// The first coarse node i >= 1 with v < g[i], or -1 (the reference's linear
// searches): from a guess taken at v's fine node, down while the node below
// also qualifies, then up to the first that does (g is increasing).
RAY_INLINE int ray_coarse_upper(const float *g, int n, const int *map, int nf, double g0, double step, float v) {
    const double f = ((double)v - g0) / step;
    int i = (f >= 0.0 && f < (double)nf) ? map[(int)f] : 1;
    i = i < 1 ? 1 : (i > n ? n : i);
    while (i > 1 && v < g[i - 1]) i--;
    while (i < n && !(v < g[i])) i++;
    return i < n ? i : -1;
}

// One segment's length at its midpoint onto the coarse derivatives: the
// eight corners of the midpoint's coarse cell (x index the upper node, as
// the reference indexes them), or with ido1d its two layers.  acc.add(v, c)
// performs du[v] = (float)(du[v] + c); indices past the variables are not
// kept (the reference never reads them).
template <class Acc>
RAY_INLINE bool ray_add_segment(const RayGrid &g, Acc &acc, float xm, float ym, float zm, double dlen) {
    const int i1 = ray_coarse_upper(g.gx, g.nxc, g.mx, g.nx, g.x0, g.df, xm);
    const int j1 = ray_coarse_upper(g.gy, g.nyc, g.my, g.ny, g.y0, g.dq, ym);
    const int k1 = ray_coarse_upper(g.gz, g.nzc, g.mz, g.nz, g.z0, g.h, zm);
    if (i1 < 0 || j1 < 0 || k1 < 0) return false;
    const float hx = g.gx[i1] - g.gx[i1 - 1];
    const float hy = g.gy[j1] - g.gy[j1 - 1];
    const float hz = g.gz[k1] - g.gz[k1 - 1];
    const double fx = (float)((xm - g.gx[i1 - 1]) / hx);
    const double fy = (float)((ym - g.gy[j1 - 1]) / hy);
    const double fz = (float)((zm - g.gz[k1 - 1]) / hz);
    const double ux = 1.0 - fx, uy = 1.0 - fy, uz = 1.0 - fz;
    const float ds[8] = {(float)(uz * uy * ux), (float)(uz * uy * fx), (float)(uz * fy * ux), (float)(uz * fy * fx),
                         (float)(fz * uy * ux), (float)(fz * uy * fx), (float)(fz * fy * ux), (float)(fz * fy * fx)};
    if (g.ido1d) {
        for (int q = 0; q < 4; q++) {
            acc.add(k1 - 1, dlen * ds[q]);
            acc.add(k1, dlen * ds[q + 4]);
        }
        return true;
    }
    const int nxyc = g.nxc * g.nyc;
    const int b0 = nxyc * (k1 - 1) + g.nxc * (j1 - 1) + i1;
    const int off[8] = {0, 1, g.nxc, g.nxc + 1, nxyc, nxyc + 1, nxyc + g.nxc, nxyc + g.nxc + 1};
    for (int q = 0; q < 8; q++) acc.add(b0 + off[q], (double)ds[q] * dlen);
    return true;
}

template <class M, class Acc>
RAY_HD void ray_trace(const RayGrid &g, const float *__restrict__ t, const RayIn &in, RayOut &out, Acc &acc) {
    out.status = kRayOk;
    out.nseg = 0;
    out.ndu = 0;
    out.first_in_cube = 0;
    out.tev = 0;
    out.dtdx = out.dtdy = out.dtdz = 0.0f;
    out.sf0 = out.sq0 = out.sr0 = 0.0;
    out.rdev = 0.0f;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long nxy = (long long)nx * ny, nxyz = nxy * nz;
    const double x0 = g.x0, y0 = g.y0, z0 = g.z0, df = g.df, dq = g.dq, h = g.h;
    const double ex = in.ex, ey = in.ey, ez = in.ez, xs = in.xs, ys = in.ys, zs = in.zs;
    const double *trow = g.trig + 4;                          // row j at trow[4*j]
    const double *tcol = g.trig + 4 * (ny + 2) + 4;           // column i at tcol[4*i]

    // the event's cell; its corner as floats (sphrayderv.c:954-987)
    const int ies = ray_to_int((ex - x0) / df), jes = ray_to_int((ey - y0) / dq), kes = ray_to_int((ez - z0) / h);
    if (ex < x0 || ies >= nx - 1 || ey < y0 || jes >= ny - 1 || ez < z0 || kes >= nz - 1) {
        out.status = kRayEventOut;
        return;
    }
    const float xis = (float)(df * ies + x0), yjs = (float)(dq * jes + y0), zks = (float)(h * kes + z0);
    // the station's nearest node and cell (sphrayderv.c:1102-1109)
    const int is = ray_to_int(round((xs - x0) / df)), js = ray_to_int(round((ys - y0) / dq)),
              ks = ray_to_int(round((zs - z0) / h));
    const int iscell = ray_to_int((xs - x0) / df), jscell = ray_to_int((ys - y0) / dq),
              kscell = ray_to_int((zs - z0) / h);
    float rsx = (float)(xs - ex), rsy = (float)(ys - ey);
    const float smag0 = sqrtf(rsx * rsx + rsy * rsy);
    rsx /= smag0;
    rsy /= smag0;

    // trilinear table time at the event and its derivatives (1322-1445)
    {
        const long long b = nxy * kes + (long long)nx * jes + ies;
        const float c[8] = {t[b], t[b + 1], t[b + nx], t[b + nx + 1], t[b + nxy], t[b + nxy + 1], t[b + nxy + nx],
                            t[b + nxy + nx + 1]};
        for (int q = 0; q < 8; q++)
            if (c[q] > 1.E9) {
                out.status = kRayUnreached;
                return;
            }
        const double fx = (ex - (double)xis) / df, fy = (ey - (double)yjs) / dq, fz = (ez - (double)zks) / h;
        const double ux = 1.0 - fx, uy = 1.0 - fy, uz = 1.0 - fz;
        const float ds[8] = {(float)(uz * uy * ux), (float)(uz * uy * fx), (float)(uz * fy * ux), (float)(uz * fy * fx),
                             (float)(fz * uy * ux), (float)(fz * uy * fx), (float)(fz * fy * ux), (float)(fz * fy * fx)};
        double tev = 0;
        for (int q = 0; q < 8; q++) tev += (double)(ds[q] * c[q]);
        out.tev = tev;
        const float wx[8] = {(float)(-uz * uy), (float)(uz * uy), (float)(-uz * fy), (float)(uz * fy),
                             (float)(-fz * uy), (float)(fz * uy), (float)(-fz * fy), (float)(fz * fy)};
        const float wy[8] = {(float)(-uz * ux), (float)(-uz * fx), (float)(uz * ux), (float)(uz * fx),
                             (float)(-fz * ux), (float)(-fz * fx), (float)(fz * ux), (float)(fz * fx)};
        const float wz[8] = {(float)(-uy * ux), (float)(-uy * fx), (float)(-fy * ux), (float)(-fy * fx),
                             (float)(uy * ux), (float)(uy * fx), (float)(fy * ux), (float)(fy * fx)};
        float tx = 0, ty = 0, tz = 0;
        for (int q = 0; q < 8; q++) {
            tx += wx[q] * c[q];
            ty += wy[q] * c[q];
            tz += wz[q] * c[q];
        }
        out.dtdx = (float)(tx / df);
        out.dtdy = (float)(ty / dq);
        out.dtdz = (float)(tz / h);
    }

    const double tolmin = h * (double)1.E-6f, tolmax = h * (double)6.0f;
    double xx = ex, yy = ey, zz = ez;
    int ie = ies, je = jes, ke = kes;
    int nseg = 0, iseg = 0, isegm1 = 0, isegm2 = 0;
    float rdevs = 0;
    int status = kRayOk;
    // the position's sines and cosines, when the last step's arctangents
    // delivered them
    double sxx = 0, cxx = 0, syy = 0, cyy = 0;
    bool hxx = false, hyy = false;
    for (;;) {
        const double zk = h * ke + z0;
        double sin_yy = syy, cos_yy = cyy;
        if (!hyy) M::sincos(yy, sin_yy, cos_yy);
        // the station's cell, or a short distance, ends the ray (1467-1575)
        const float dxs = (float)(xx - xs), dys = (float)(yy - ys), dzs = (float)(zz - zs);
        const float rs = (float)(kREarth - zz);
        const float dqs = rs * dys;
        const float dfs = (float)((double)rs * sin_yy * dxs);
        const double dl_st = sqrtf(dfs * dfs + dqs * dqs + dzs * dzs);
        if ((ie == iscell && je == jscell && ke == kscell) || dl_st < h / (double)1000.f) {
            const float xm = (float)(xx - dxs * 0.5f), ym = (float)(yy - dys * 0.5f), zm = (float)(zz - dzs * 0.5f);
            if (!ray_add_segment(g, acc, xm, ym, zm, dl_st)) status = kRayCoarse;
            break;
        }
        if (nseg >= g.max_segments) {
            status = kRayLimit;
            break;
        }
        const long long b = nxy * ke + (long long)nx * je + ie;
        if (b < 0 || b + nxy + nx + 1 >= nxyz) {
            status = kRayEdge;
            break;
        }
        const float c[8] = {t[b], t[b + 1], t[b + nx], t[b + nx + 1], t[b + nxy], t[b + nxy + 1], t[b + nxy + nx],
                            t[b + nxy + nx + 1]};
        // the cell's row and column sines and cosines
        const double *rw = trow + 4 * je, *cl = tcol + 4 * ie;
        // -grad T from the corners (1786-1815)
        const double r1 = kREarth - zk, r2 = r1 - h;
        double gr[3];
        gr[0] = ((double)(c[1] - c[0]) / (r1 * rw[0]) + (double)(c[3] - c[2]) / (r1 * rw[2]) +
                 (double)(c[5] - c[4]) / (r2 * rw[0]) + (double)(c[7] - c[6]) / (r2 * rw[2])) / (4 * df);
        gr[1] = ((double)(c[2] + c[3] - c[0] - c[1]) / r1 + (double)(c[6] + c[7] - c[4] - c[5]) / r2) / (4 * dq);
        gr[2] = (double)(c[4] + c[5] + c[6] + c[7] - c[0] - c[1] - c[2] - c[3]) / (4 * h);
        gr[0] = -gr[0];
        gr[1] = -gr[1];
        // position and step direction (1817-1863)
        double sinf_ = sxx, cosf_ = cxx;
        if (!hxx) M::sincos(xx, sinf_, cosf_);
        const double cosq0 = cos_yy, sinq0 = sin_yy;
        const double ro = kREarth - zz;
        const double xo = ro * sinq0 * cosf_, yo = ro * sinq0 * sinf_, zo = ro * cosq0;
        double sx, sy, sz, sf = 0, sq = 0, sr = 0;
        const bool in_cube = ie >= is - 2 && ie < is + 2 && je >= js - 2 && je < js + 2 && ke >= ks - 2 && ke < ks + 2;
        if (in_cube) {
            const double dx = in.x2 - xo, dy = in.y2 - yo, dz = in.z2 - zo;
            const double dl = sqrt(dx * dx + dy * dy + dz * dz);
            sx = dx / dl;
            sy = dy / dl;
            sz = dz / dl;
        } else {
            const double gm = sqrt(gr[0] * gr[0] + gr[1] * gr[1] + gr[2] * gr[2]);
            sf = gr[0] / gm;
            sq = gr[1] / gm;
            sr = gr[2] / gm;
            sx = -sinf_ * sf + cosq0 * cosf_ * sq + sinq0 * cosf_ * sr;
            sy = cosf_ * sf + cosq0 * sinf_ * sq + sinq0 * sinf_ * sr;
            sz = -sinq0 * sq + cosq0 * sr;
            const float smag = sqrtf((float)(sx * sx + sy * sy + sz * sz));
            sx /= smag;
            sy /= smag;
            sz /= smag;
        }
        if (nseg == 0) {
            if (in_cube) {
                out.first_in_cube = 1;
            } else {
                out.sf0 = sf;
                out.sq0 = sq;
                out.sr0 = sr;
            }
        }
        // distances to the cell's surfaces of constant radius, colatitude,
        // longitude (1875-1955)
        double dd[3];
        dd[2] = -1e20;
        if (gr[2] != 0) {
            const double rr = gr[2] > 0 ? kREarth - zk : kREarth - zk - h;
            const double bq = xo * sx + yo * sy + zo * sz;
            const double cq = xo * xo + yo * yo + zo * zo - rr * rr;
            const double quad = bq * bq - cq;
            if (quad >= 0) {
                const double sq_ = sqrt(quad);
                dd[2] = ray_pmin((float)(-bq + sq_), (float)(-bq - sq_));
            }
        }
        dd[1] = -1e20;
        if (gr[1] != 0) {
            const double sinq = gr[1] < 0 ? rw[0] : rw[2], cosq = gr[1] < 0 ? rw[1] : rw[3];
            double aq, bq, cq;
            if (fabs(cosq) > fabs(sinq)) {
                const double tanq = sinq / cosq, tansq = tanq * tanq;
                aq = sx * sx + sy * sy - sz * sz * tansq;
                bq = xo * sx + yo * sy - zo * sz * tansq;
                cq = xo * xo + yo * yo - zo * zo * tansq;
            } else {
                const double ctanq = cosq / sinq, ctansq = ctanq * ctanq;
                aq = (sx * sx + sy * sy) * ctansq - sz * sz;
                bq = (xo * sx + yo * sy) * ctansq - zo * sz;
                cq = (xo * xo + yo * yo) * ctansq - zo * zo;
            }
            const double quad = bq * bq - aq * cq;
            if (quad >= 0) {
                const double sq_ = sqrt(quad);
                dd[1] = ray_pmin((float)((-bq + sq_) / aq), (float)((-bq - sq_) / aq));
            }
        }
        dd[0] = -1e20;
        if (gr[0] != 0) {
            const double sinf = gr[0] < 0 ? cl[0] : cl[2], cosf = gr[0] < 0 ? cl[1] : cl[3];
            float d1;
            if (fabs(cosf) > fabs(sinf)) {
                const double tanf = sinf / cosf;
                d1 = (float)((xo * tanf - yo) / (sy - sx * tanf));
            } else {
                const double ctanf = cosf / sinf;
                d1 = (float)((yo * ctanf - xo) / (sx - sy * ctanf));
            }
            if (d1 > 0) dd[0] = d1;
        }
        // no step back into the cell just left, none closing a four-cell
        // loop (1957-2026)
        const int sum3 = iseg + isegm1 + isegm2;
        const int code[3] = {1, 20, 300};
        for (int a = 0; a < 3; a++) {
            // the step's sign along the axis: -gr for longitude and
            // colatitude, gr for radius; no sign (0 or NaN): no step
            const double s = a == 2 ? -gr[2] : gr[a];
            if (s < 0 || s > 0) {
                const int want = s < 0 ? code[a] : -code[a];
                if (dd[a] < h / 1000 && (iseg == want || (sum3 == want && nseg >= 3))) {
                    gr[a] = 0;
                    dd[a] = -1e20;
                }
            } else {
                dd[a] = -1e20;
            }
        }
        double d;
        int md;
        ray_pick(dd, tolmin, tolmax, d, md);
        // the step (2031-2039)
        const double xold = xx, yold = yy, zold = zz;
        const double xn = xo + d * sx, yn = yo + d * sy, zn = zo + d * sz;
        const double rxy = sqrt(xn * xn + yn * yn);
        xx = M::atan2_sc(yn, xn, sxx, cxx, hxx);
        yy = M::atan2_sc(rxy, zn, syy, cyy, hyy);
        zz = kREarth - sqrt(xn * xn + yn * yn + zn * zn);
        if (xx != xx || yy != yy || zz != zz) {
            status = kRayEdge;
            break;
        }
        nseg++;
        isegm2 = isegm1;
        isegm1 = iseg;
        // the next cell (2043-2099)
        bool left = false;
        if (md == 0) {
            je = ray_to_int((yy - y0) / dq);
            ke = ray_to_int((zz - z0) / h);
            if (gr[0] < 0) {
                if (ie < 0) left = true; else { ie--; iseg = -1; }
            } else {
                if (ie >= nx - 1) left = true; else { ie++; iseg = 1; }
            }
        } else if (md == 1) {
            if (gr[1] < 0) {
                if (je < 0) left = true; else { je--; iseg = -20; }
            } else {
                if (je >= ny - 1) left = true; else { je++; iseg = 20; }
            }
            if (!left) {
                ie = ray_to_int((xx - x0) / df);
                ke = ray_to_int((zz - z0) / h);
            }
        } else {
            if (gr[2] > 0) {
                if (ke < 0) left = true; else { ke--; iseg = -300; }
            } else {
                if (ke >= nz - 1) left = true; else { ke++; iseg = 300; }
            }
            if (!left) {
                ie = ray_to_int((xx - x0) / df);
                je = ray_to_int((yy - y0) / dq);
            }
        }
        if (left || xx < x0 || ie >= nx || yy < y0 || je >= ny || zz < z0 || ke >= nz) {
            status = kRayLeft;
            break;
        }
        // the segment's derivatives at its midpoint (2101-2202)
        const float xm = (float)((xx + xold) * 0.5), ym = (float)((yy + yold) * 0.5), zm = (float)((zz + zold) * 0.5);
        if (!ray_add_segment(g, acc, xm, ym, zm, d)) {
            status = kRayCoarse;
            break;
        }
        // deviation from the source-station plane (2238-2246)
        const float rx = (float)(xm - ex), ry = (float)(ym - ey);
        const float rmagsq = rx * rx + ry * ry, rds = rsx * rx + rsy * ry;
        rdevs += sqrtf(rmagsq - rds * rds);
    }
    out.status = status;
    out.nseg = nseg;
    out.rdev = rdevs / nseg;
}

}  // namespace fdray
