/*
 * ray_oracle.c -- CPU restatement of sphrayderv's ray tracing (event
 * interpolation src/sphrayderv.c:1306-1445, ray loop 1466-2250, dfind and
 * pmin of lib/sphrayderv).  Test checker only; the header states how it is
 * pinned.
 *
 * Arithmetic follows the reference's declared types and evaluation order
 * (x86-64: float expressions in float, double in double, no contraction --
 * built with -ffp-contract=off): positions, directions, intersection
 * distances and the gradient sums in double; table differences, the
 * trilinear weights, segment midpoints, the path-length derivatives and the
 * event derivatives in float where the reference stores them as float.
 */
#include "ray_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* sphrayderv.c:96-97: degrad is a float literal held in a double, rearth a float. */
static const double kDegRad = 0.017453292f;
static const float kREarth = 6371.0f;

/* (int) conversion as x86-64 performs it where the reference converts
 * values that can be out of int range (huge depths of a diverging step):
 * INT_MIN for NaN and out-of-range values. */
static int to_int(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
    return (int)v;
}

/* pmin (lib/sphrayderv/pmin.c): the smaller non-negative of two distances,
 * or -1e20 when neither is non-negative. */
static double pmin2(double d1, double d2) {
    if (d1 >= 0) return (d2 >= 0 && d2 <= d1) ? d2 : d1;
    return d2 >= 0 ? d2 : -1e20;
}

/* dfind (lib/sphrayderv/dfind.c): order the three candidate distances by
 * magnitude (ties keep the lower axis first, as its comparisons do) and take
 * the smallest unless it is below tolmin; then the middle one unless it too
 * is below tolmin (the middle one is only taken when it is below tolmax);
 * else the largest when below tolmax, else the middle. */
static void pick_step(const double dd[3], double tolmin, double tolmax, double *d, int *md) {
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
    *d = dd[mn];
    *md = mn;
    if (dd[mn] < tolmin) {
        if (dd[mid] < tolmin) {
            if (dd[mx] < tolmax) { *d = dd[mx]; *md = mx; }
            else { *d = dd[mid]; *md = mid; }
        } else if (dd[mid] < tolmax) {
            *d = dd[mid];
            *md = mid;
        }
    }
}

/* First coarse node above v (the reference's linear searches,
 * sphrayderv.c:1492-1528, 2107-2140), or -1. */
static int coarse_upper(const float *g, int n, float v) {
    for (int i = 1; i < n; i++)
        if (v < g[i]) return i;
    return -1;
}

/* Adds one segment of length dlen with midpoint (xm, ym, zm) to the coarse
 * derivatives (sphrayderv.c:1530-1573 and 2142-2202).  The x index of the
 * eight corners is the upper node i1, not i1 - 1, as the reference indexes
 * them.  Returns 0, or -1 when the midpoint is beyond the coarse grid. */
static int add_segment(const oracle_ray_grid *g, float *du, float xm, float ym, float zm, double dlen) {
    const int i1 = coarse_upper(g->gx, g->nxc, xm);
    const int j1 = coarse_upper(g->gy, g->nyc, ym);
    const int k1 = coarse_upper(g->gz, g->nzc, zm);
    if (i1 < 0 || j1 < 0 || k1 < 0) return -1;
    const float hx = g->gx[i1] - g->gx[i1 - 1];
    const float hy = g->gy[j1] - g->gy[j1 - 1];
    const float hz = g->gz[k1] - g->gz[k1 - 1];
    /* float quotients, held in the reference's double fx/fy/fz */
    const double fx = (float)((xm - g->gx[i1 - 1]) / hx);
    const double fy = (float)((ym - g->gy[j1 - 1]) / hy);
    const double fz = (float)((zm - g->gz[k1 - 1]) / hz);
    float ds[8];
    ds[0] = (float)((1.0 - fz) * (1.0 - fy) * (1.0 - fx));
    ds[1] = (float)((1.0 - fz) * (1.0 - fy) * fx);
    ds[2] = (float)((1.0 - fz) * fy * (1.0 - fx));
    ds[3] = (float)((1.0 - fz) * fy * fx);
    ds[4] = (float)(fz * (1.0 - fy) * (1.0 - fx));
    ds[5] = (float)(fz * (1.0 - fy) * fx);
    ds[6] = (float)(fz * fy * (1.0 - fx));
    ds[7] = (float)(fz * fy * fx);
    if (g->ido1d) {
        for (int q = 0; q < 4; q++) {
            du[k1 - 1] = (float)((double)du[k1 - 1] + dlen * ds[q]);
            du[k1] = (float)((double)du[k1] + dlen * ds[q + 4]);
        }
        return 0;
    }
    const int nxyc = g->nxc * g->nyc;
    const int b0 = nxyc * (k1 - 1) + g->nxc * (j1 - 1) + i1;
    const int cj = g->nxc, ck = nxyc;
    const int pt[8] = {b0, b0 + 1, b0 + cj, b0 + cj + 1, b0 + ck, b0 + ck + 1, b0 + ck + cj, b0 + ck + cj + 1};
    for (int q = 0; q < 8; q++) du[pt[q]] = (float)((double)du[pt[q]] + (double)ds[q] * dlen);
    return 0;
}

int oracle_ray_trace(const oracle_ray_grid *g, const float *t, const oracle_ray_in *in, oracle_ray_out *out,
                     float *du, int *idx, float *val, int cap) {
    memset(out, 0, sizeof(*out));
    out->ai = out->az = NAN;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const long long nxy = (long long)nx * ny, nxyz = nxy * nz;
    const double x0 = g->x0, y0 = g->y0, z0 = g->z0, df = g->df, dq = g->dq, h = g->h;
    const double ex = in->ex, ey = in->ey, ez = in->ez, xs = in->xs, ys = in->ys, zs = in->zs;
    const int maxvarc = g->ido1d ? g->nzc : g->nxc * g->nyc * g->nzc;
    memset(du, 0, sizeof(float) * (size_t)(maxvarc + 1));

    /* the event's cell (sphrayderv.c:954-985); its corner 1 as floats */
    const int ies = to_int((ex - x0) / df), jes = to_int((ey - y0) / dq), kes = to_int((ez - z0) / h);
    if (ex < x0 || ies >= nx - 1 || ey < y0 || jes >= ny - 1 || ez < z0 || kes >= nz - 1) {
        out->status = ORACLE_RAY_EVENT_OUT;
        return 0;
    }
    const float xis = (float)(df * ies + x0), yjs = (float)(dq * jes + y0), zks = (float)(h * kes + z0);
    /* the station's nearest node and its cell (sphrayderv.c:1102-1109) */
    const int is = to_int(round((xs - x0) / df)), js = to_int(round((ys - y0) / dq)), ks = to_int(round((zs - z0) / h));
    const int iscell = to_int((xs - x0) / df), jscell = to_int((ys - y0) / dq), kscell = to_int((zs - z0) / h);
    /* unit vector event -> station for the deviation statistic (1230-1300) */
    float rsx = (float)(xs - ex), rsy = (float)(ys - ey);
    const float smag0 = sqrtf(rsx * rsx + rsy * rsy);
    rsx /= smag0;
    rsy /= smag0;

    /* the table at the event by trilinear interpolation, and its
     * derivatives (sphrayderv.c:1322-1445) */
    {
        const long long b = nxy * kes + (long long)nx * jes + ies;
        const long long pt[8] = {b, b + 1, b + nx, b + nx + 1, b + nxy, b + nxy + 1, b + nxy + nx, b + nxy + nx + 1};
        float c[8];
        for (int q = 0; q < 8; q++) c[q] = t[pt[q]];
        for (int q = 0; q < 8; q++)
            if (c[q] > 1.E9) {
                out->status = ORACLE_RAY_UNREACHED;
                return 0;
            }
        const double fx = (ex - (double)xis) / df, fy = (ey - (double)yjs) / dq, fz = (ez - (double)zks) / h;
        const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
        const float ds[8] = {(float)(gz * gy * gx), (float)(gz * gy * fx), (float)(gz * fy * gx), (float)(gz * fy * fx),
                             (float)(fz * gy * gx), (float)(fz * gy * fx), (float)(fz * fy * gx), (float)(fz * fy * fx)};
        double tev = 0;
        for (int q = 0; q < 8; q++) tev += (double)(ds[q] * c[q]);
        out->tev = tev;
        const float wx[8] = {(float)(-gz * gy), (float)(gz * gy), (float)(-gz * fy), (float)(gz * fy),
                             (float)(-fz * gy), (float)(fz * gy), (float)(-fz * fy), (float)(fz * fy)};
        const float wy[8] = {(float)(-gz * gx), (float)(-gz * fx), (float)(gz * gx), (float)(gz * fx),
                             (float)(-fz * gx), (float)(-fz * fx), (float)(fz * gx), (float)(fz * fx)};
        const float wz[8] = {(float)(-gy * gx), (float)(-gy * fx), (float)(-fy * gx), (float)(-fy * fx),
                             (float)(gy * gx), (float)(gy * fx), (float)(fy * gx), (float)(fy * fx)};
        float tx = 0, ty = 0, tz = 0;
        for (int q = 0; q < 8; q++) {
            tx += wx[q] * c[q];
            ty += wy[q] * c[q];
            tz += wz[q] * c[q];
        }
        out->dtdx = (float)(tx / df);
        out->dtdy = (float)(ty / dq);
        out->dtdz = (float)(tz / h);
    }

    const double tolmin = h * 1.E-6f, tolmax = h * 6.0f;
    double xx = ex, yy = ey, zz = ez;
    int ie = ies, je = jes, ke = kes;
    int nseg = 0, iseg = 0, isegm1 = 0, isegm2 = 0;
    double sf = 0, sq = 0, sr = 0;
    float rdevs = 0;
    int status = ORACLE_RAY_OK;
    for (;;) {
        /* distance to the station; the station's cell or a short distance
         * ends the ray with the segment to the station (1467-1575) */
        const double xi = df * ie + x0, yj = dq * je + y0, zk = h * ke + z0;
        const float dxs = (float)(xx - xs), dys = (float)(yy - ys), dzs = (float)(zz - zs);
        const float rs = (float)(kREarth - zz);
        const float dqs = rs * dys;
        const float dfs = (float)((double)rs * sin(yy) * dxs);
        const double dl_st = sqrtf(dfs * dfs + dqs * dqs + dzs * dzs);
        if ((ie == iscell && je == jscell && ke == kscell) || dl_st < h / 1000.f) {
            const float xm = (float)(xx - dxs * 0.5f), ym = (float)(yy - dys * 0.5f), zm = (float)(zz - dzs * 0.5f);
            if (add_segment(g, du, xm, ym, zm, dl_st)) status = ORACLE_RAY_COARSE;
            break;
        }
        if (nseg >= g->max_segments) {
            status = ORACLE_RAY_LIMIT;
            break;
        }
        /* the cell's corners (1756-1769) */
        const long long b = nxy * ke + (long long)nx * je + ie;
        const long long pt[8] = {b, b + 1, b + nx, b + nx + 1, b + nxy, b + nxy + 1, b + nxy + nx, b + nxy + nx + 1};
        int edge = 0;
        for (int q = 0; q < 8; q++) edge |= pt[q] < 0 || pt[q] >= nxyz;
        if (edge) {
            status = ORACLE_RAY_EDGE;
            break;
        }
        float c[8];
        for (int q = 0; q < 8; q++) c[q] = t[pt[q]];
        /* -grad T in (longitude, colatitude, radius) from the corners
         * (1786-1815): corner differences in float, the rest in double */
        const double r1 = kREarth - zk, r2 = r1 - h, q1 = yj, q2 = q1 + dq;
        double gr[3];
        gr[0] = ((double)(c[1] - c[0]) / (r1 * sin(q1)) + (double)(c[3] - c[2]) / (r1 * sin(q2)) +
                 (double)(c[5] - c[4]) / (r2 * sin(q1)) + (double)(c[7] - c[6]) / (r2 * sin(q2))) / (4 * df);
        gr[1] = ((double)(c[2] + c[3] - c[0] - c[1]) / r1 + (double)(c[6] + c[7] - c[4] - c[5]) / r2) / (4 * dq);
        gr[2] = (double)(c[4] + c[5] + c[6] + c[7] - c[0] - c[1] - c[2] - c[3]) / (4 * h);
        gr[0] = -gr[0];
        gr[1] = -gr[1];
        /* position in Cartesian coordinates and the step direction: straight
         * to the station within two cells of its node, else down the
         * gradient (1817-1863) */
        const double sinf_ = sin(xx), cosf_ = cos(xx), sinq0 = sin(yy), cosq0 = cos(yy);
        const double ro = kREarth - zz;
        const double xo = ro * sinq0 * cosf_, yo = ro * sinq0 * sinf_, zo = ro * cosq0;
        double sx, sy, sz;
        const int in_cube = ie >= is - 2 && ie < is + 2 && je >= js - 2 && je < js + 2 && ke >= ks - 2 && ke < ks + 2;
        if (in_cube) {
            const double rst = kREarth - zs;
            const float x2 = (float)(rst * sin(ys) * cos(xs)), y2 = (float)(rst * sin(ys) * sin(xs)),
                        z2 = (float)(rst * cos(ys));
            const double dx = x2 - xo, dy = y2 - yo, dz = z2 - zo;
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
            /* take-off angles from the first gradient direction (1865-1873);
             * a first segment in the station cube leaves them to the
             * reference's previous ray */
            if (in_cube) {
                out->first_in_cube = 1;
            } else {
                const double sfq = sqrt(sf * sf + sq * sq);
                out->az = (float)(atan2f((float)sf, (float)-sq) / kDegRad);
                out->ai = (float)(atan2f((float)sfq, (float)-sr) / kDegRad);
            }
        }
        /* distances along the direction to the cell's bounding surfaces of
         * constant radius, colatitude and longitude (1875-1955) */
        double dd[3];
        dd[2] = -1e20;
        if (gr[2] != 0) {
            const double rr = gr[2] > 0 ? kREarth - zk : kREarth - zk - h;
            const double bq = xo * sx + yo * sy + zo * sz;
            const double cq = xo * xo + yo * yo + zo * zo - rr * rr;
            const double quad = bq * bq - cq;
            if (quad >= 0) {
                const float d1 = (float)(-bq + sqrt(quad)), d2 = (float)(-bq - sqrt(quad));
                dd[2] = pmin2(d1, d2);
            }
        }
        dd[1] = -1e20;
        if (gr[1] != 0) {
            const double qb = gr[1] < 0 ? yj : yj + dq;
            const double sinq = sin(qb), cosq = cos(qb);
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
                const float d1 = (float)((-bq + sqrt(quad)) / aq), d2 = (float)((-bq - sqrt(quad)) / aq);
                dd[1] = pmin2(d1, d2);
            }
        }
        dd[0] = -1e20;
        if (gr[0] != 0) {
            const double fb = gr[0] < 0 ? xi : xi + df;
            const double sinf = sin(fb), cosf = cos(fb);
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
        /* no step back into the cell just left, and none that closes a loop
         * over four cells (1957-2026) */
        const int sum3 = iseg + isegm1 + isegm2;
        const int code[3] = {1, 20, 300};
        for (int a = 0; a < 3; a++) {
            /* the sign of the step along this axis: longitude and colatitude
             * steps follow -gr, radius steps follow gr */
            const double s = a == 2 ? -gr[2] : gr[a];
            if (s < 0) {
                if (dd[a] < h / 1000 && (iseg == code[a] || (sum3 == code[a] && nseg >= 3))) {
                    gr[a] = 0;
                    dd[a] = -1e20;
                }
            } else if (s > 0) {
                if (dd[a] < h / 1000 && (iseg == -code[a] || (sum3 == -code[a] && nseg >= 3))) {
                    gr[a] = 0;
                    dd[a] = -1e20;
                }
            } else {
                dd[a] = -1e20;
            }
        }
        double d;
        int md;
        pick_step(dd, tolmin, tolmax, &d, &md);
        /* the step, back to spherical coordinates (2031-2039) */
        const double xold = xx, yold = yy, zold = zz;
        const double xn = xo + d * sx, yn = yo + d * sy, zn = zo + d * sz;
        xx = atan2(yn, xn);
        yy = atan2(sqrt(xn * xn + yn * yn), zn);
        zz = kREarth - sqrt(xn * xn + yn * yn + zn * zn);
        if (xx != xx || yy != yy || zz != zz) {
            status = ORACLE_RAY_EDGE;
            break;
        }
        nseg++;
        isegm2 = isegm1;
        isegm1 = iseg;
        /* the next cell: across the face the step ended on (2043-2099) */
        int left = 0;
        if (md == 0) {
            je = to_int((yy - y0) / dq);
            ke = to_int((zz - z0) / h);
            if (gr[0] < 0) {
                if (ie < 0) left = 1; else { ie--; iseg = -1; }
            } else {
                if (ie >= nx - 1) left = 1; else { ie++; iseg = 1; }
            }
        } else if (md == 1) {
            if (gr[1] < 0) {
                if (je < 0) left = 1; else { je--; iseg = -20; }
            } else {
                if (je >= ny - 1) left = 1; else { je++; iseg = 20; }
            }
            if (!left) {
                ie = to_int((xx - x0) / df);
                ke = to_int((zz - z0) / h);
            }
        } else {
            if (gr[2] > 0) {
                if (ke < 0) left = 1; else { ke--; iseg = -300; }
            } else {
                if (ke >= nz - 1) left = 1; else { ke++; iseg = 300; }
            }
            if (!left) {
                ie = to_int((xx - x0) / df);
                je = to_int((yy - y0) / dq);
            }
        }
        if (left || xx < x0 || ie >= nx || yy < y0 || je >= ny || zz < z0 || ke >= nz) {
            status = ORACLE_RAY_LEFT;
            break;
        }
        /* the segment's derivatives at its midpoint (2101-2202) */
        const float xm = (float)((xx + xold) * 0.5), ym = (float)((yy + yold) * 0.5), zm = (float)((zz + zold) * 0.5);
        if (add_segment(g, du, xm, ym, zm, d)) {
            status = ORACLE_RAY_COARSE;
            break;
        }
        /* deviation from the source-station plane (2238-2246) */
        const float rx = (float)(xm - ex), ry = (float)(ym - ey);
        const float rmagsq = rx * rx + ry * ry, rds = rsx * rx + rsy * ry;
        rdevs += sqrtf(rmagsq - rds * rds);
    }
    out->status = status;
    out->nseg = nseg;
    out->rdev = rdevs / nseg;
    if (status != ORACLE_RAY_OK) return 0;
    /* the nonzero derivatives in index order (1650-1663) */
    int n = 0;
    for (int j = 0; j < maxvarc; j++)
        if (du[j] != 0) {
            if (n < cap) {
                idx[n] = j;
                val[n] = du[j];
            }
            n++;
        }
    out->ndu = n;
    return n <= cap ? 0 : -1;
}

typedef struct {
    const oracle_ray_grid *g;
    const float *const *tables;
    const oracle_ray_in *rays;
    oracle_ray_out *out;
    int *idx;
    float *val;
    int cap, n, stride, first, overflow;
} trace_job;

static void *trace_worker(void *p) {
    trace_job *J = (trace_job *)p;
    const int m = (J->g->ido1d ? J->g->nzc : J->g->nxc * J->g->nyc * J->g->nzc) + 1;
    float *du = (float *)malloc(sizeof(float) * (size_t)m);
    if (!du) {
        J->overflow = -1;
        return NULL;
    }
    for (int r = J->first; r < J->n; r += J->stride)
        if (oracle_ray_trace(J->g, J->tables[r], &J->rays[r], &J->out[r], du, J->idx + (size_t)r * J->cap,
                             J->val + (size_t)r * J->cap, J->cap))
            J->overflow++;
    free(du);
    return NULL;
}

int oracle_ray_trace_many(const oracle_ray_grid *g, const float *const *tables, const oracle_ray_in *rays, int n,
                          oracle_ray_out *out, int *idx, float *val, int cap_per_ray, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    trace_job jobs[256];
    for (int k = 0; k < nthreads; k++) {
        jobs[k] = (trace_job){g, tables, rays, out, idx, val, cap_per_ray, n, nthreads, k, 0};
        if (k > 0) pthread_create(&th[k], NULL, trace_worker, &jobs[k]);
    }
    trace_worker(&jobs[0]);
    int over = jobs[0].overflow;
    for (int k = 1; k < nthreads; k++) {
        pthread_join(th[k], NULL);
        over += jobs[k].overflow;
    }
    return over;
}
