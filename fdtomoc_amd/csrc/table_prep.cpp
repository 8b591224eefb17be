// table_prep.cpp -- see table_prep.h.  Compiled with -ffp-contract=off.
#include "table_prep.h"

#include <math.h>
#include <string.h>

namespace fdprep {

namespace {
// The reference's constants (sphfd.c:144-150).  degrad is used there only as
// `x *= degrad`, i.e. x * (PI/180.0).
constexpr double kPi = 3.141592654;
constexpr double kHpi = 1.570796327;
constexpr double kRearth = 6371.0;
constexpr double kDegrad = kPi / 180.0;

// WGS84 geodetic latitude + ellipsoidal depth -> geocentric latitude and
// radius (sphfd.c:5836-5893).
double glath(double hlat, double h, double *r) {
    const double ap = 6378137.0;
    const double esq = 0.00669437978616733379;
    const double hm = -h * 1000.0;
    const double sinxl = sin(hlat);
    const double anu = ap / sqrt(1.0 - esq * sinxl * sinxl);
    const double xp = (anu + hm) * cos(hlat);
    const double zp = ((1.0 - esq) * anu + hm) * sinxl;
    *r = sqrt(xp * xp + zp * zp) / 1000.0;
    return atan2(zp, xp);
}
}  // namespace

bool skip_safe(const ParValues &p, const Prepared &g, double vmin, double vmax) {
    if (!(vmin > 0.0) || !(vmax >= vmin) || !(vmax < 1.0e30)) return false;
    const double smin = 1.0 / vmax, smax = 1.0 / vmin;
    const double rbot = g.z0r - g.h * p.nz;
    double sinmin = 1.0;
    for (double s : g.sinq) sinmin = fabs(s) < sinmin ? fabs(s) : sinmin;
    double dmin = fabs(g.h);
    dmin = fmin(dmin, fabs(rbot * g.dq));
    dmin = fmin(dmin, fabs(rbot * sinmin * g.df));
    if (!(dmin > 0.0)) return false;
    const double ext = fabs(g.h) * sqrt((double)p.nx * p.nx + (double)p.ny * p.ny + (double)p.nz * p.nz);
    // bound on any travel time: twice the slowest straight crossing of the grid
    const double tmax = 2.0 * smax * ext + 1.0;
    const double ulp = ldexp(1.0, ilogb(tmax) - 23);   // float spacing at tmax
    // fhead of a non-changing store is below ulp / (2 d s); require 8x margin to 1e-3
    return ulp / (dmin * smin) < 2.5e-4;
}

long long max_face(int nx, int ny, int nz) {
    long long a = (long long)nx * ny, b = (long long)ny * nz, c = (long long)nx * nz;
    long long m = a > b ? a : b;
    return m > c ? m : c;
}

int prepare_geometry(const ParValues &p, Prepared &o) {
    double h = p.h, dq = p.dq, df = p.df, x0 = p.x0, y0 = p.y0, z0 = p.z0;
    if (p.invh == 1) h = 1. / h;
    if (p.invq == 1 && dq != 0) dq = 1. / dq;
    if (p.invf == 1 && df != 0) df = 1. / df;
    if (p.invx0 == 1 && x0 != 0) x0 = 1. / x0;
    if (p.invy0 == 1 && y0 != 0) y0 = 1. / y0;
    if (p.invz0 == 1 && z0 != 0) z0 = 1. / z0;
    o.x0_deg = x0;
    o.y0_deg = y0;
    o.z0_km = z0;
    double z0r;
    y0 *= kDegrad;
    y0 = kHpi - glath(y0, z0, &z0r);
    x0 *= kDegrad;
    dq *= kDegrad;
    df *= kDegrad;
    if (dq == 0) dq = h / kRearth;
    if (df == 0) df = fabs(h / (kRearth * sin(y0)));
    o.h = h; o.dq = dq; o.df = df; o.x0 = x0; o.y0 = y0; o.z0r = z0r;
    int xs = 0, ys = 0, zs = 0;
    if (p.srctype == 1) {
        if (p.floatsrc == 0) {
            xs = p.xs; ys = p.ys; zs = p.zs;
            o.fxss = x0 + df * (xs);
            o.fyss = y0 + dq * (ys);
            o.fzss = z0r - h * (zs);
        } else {
            double fxs = p.fxs, fys = p.fys, fzs = p.fzs, rcent;
            fxs *= kDegrad;
            fys *= kDegrad;
            fys = kHpi - glath(fys, fzs, &rcent);
            o.fxss = fxs; o.fyss = fys; o.fzss = rcent;
            fxs = (fxs - x0) / df;
            fys = (fys - y0) / dq;
            fzs = (z0r - rcent) / h;
            xs = (int)(fxs + 0.5);
            ys = (int)(fys + 0.5);
            zs = (int)(fzs + 0.5);
        }
    } else if (p.srctype != 2 && p.srctype != 3) {
        return kErrSrctype;
    }
    o.xs = xs; o.ys = ys; o.zs = zs;
    o.maxrad = (p.maxoff > 0.) ? (int)(p.maxoff / h + 1) : 99999999;
    o.sinq.resize((size_t)p.ny + 2);
    for (int j = -1; j <= p.ny; j++) o.sinq[(size_t)(j + 1)] = sin(y0 + dq * (j));
    if (p.srctype == 1) {
        if (xs < 0 || ys < 0 || zs < 0 || xs > p.nx - 1 || ys > p.ny - 1 || zs > p.nz - 1) return kErrSourceOutside;
        o.near_edge = xs < 2 || ys < 2 || zs < 2 || xs > p.nx - 3 || ys > p.ny - 3 || zs > p.nz - 3;
    }
    return kOk;
}

int prepare_source(const ParValues &p, Prepared &o, const float *preset, const float *boxvals, float *grid) {
    const int nx = p.nx, ny = p.ny, nz = p.nz, NC = p.ncube;
    const long long nxy = (long long)nx * ny;
    int x1, x2, y1, y2, z1, z2, dx1 = 1, dx2 = 1, dy1 = 1, dy2 = 1, dz1 = 1, dz2 = 1;
    o.init_idx.clear();
    o.init_dist.clear();
    o.init_val.clear();
    o.srcwall = p.srcwall;
    if (p.srctype == 1) {
        const int xs = o.xs, ys = o.ys, zs = o.zs;
        o.src_index = (long long)zs * nxy + (long long)ys * nx + xs;
        // sphfd.c:727-749: exact straight-ray distances in the source cube
        const double x1p = o.fzss * cos(o.fyss);
        const double y1p = o.fzss * sin(o.fyss) * cos(o.fxss);
        const double z1p = o.fzss * sin(o.fyss) * sin(o.fxss);
        int ii = 0;
        for (int zz = zs - NC; zz <= zs + NC; zz++) {
            if (zz < 0 || zz >= nz) continue;
            for (int yy = ys - NC; yy <= ys + NC; yy++) {
                if (yy < 0 || yy >= ny) continue;
                for (int xx = xs - NC; xx <= xs + NC; xx++) {
                    if (xx < 0 || xx >= nx) continue;
                    o.init_idx.push_back((int)((long long)zz * nxy + (long long)yy * nx + xx));
                    if (p.fill == 0) {
                        const double fx1 = o.x0 + o.df * (xx), qy1 = o.y0 + o.dq * (yy), rz1 = o.z0r - o.h * (zz);
                        const double x2p = rz1 * cos(qy1);
                        const double y2p = rz1 * sin(qy1) * cos(fx1);
                        const double z2p = rz1 * sin(qy1) * sin(fx1);
                        const double d = sqrt((x1p - x2p) * (x1p - x2p) + (y1p - y2p) * (y1p - y2p) +
                                              (z1p - z2p) * (z1p - z2p));
                        o.init_dist.push_back(d);
                    } else {
                        if (!boxvals) return kErrInput;
                        o.init_val.push_back(boxvals[ii++]);
                    }
                }
            }
        }
        o.init_mode = p.fill == 0 ? 0 : 2;
        // sphfd.c:789-825
        o.radius0 = NC;
        if (xs > NC) x1 = xs - (NC + 1); else { x1 = -1; dx1 = 0; }
        if (xs < nx - (NC + 1)) x2 = xs + (NC + 1); else { x2 = nx; dx2 = 0; }
        if (ys > NC) y1 = ys - (NC + 1); else { y1 = -1; dy1 = 0; }
        if (ys < ny - (NC + 1)) y2 = ys + (NC + 1); else { y2 = ny; dy2 = 0; }
        if (zs > NC) z1 = zs - (NC + 1); else { z1 = -1; dz1 = 0; }
        if (zs < nz - (NC + 1)) z2 = zs + (NC + 1); else { z2 = nz; dz2 = 0; }
    } else {
        if (!preset || !grid) return kErrInput;
        const long long n = nxy * nz;
        if (p.srctype == 2) {
            // sphfd.c:829-888: six walls of nwall floats, each applied if its
            // first value exceeds -1e-20
            const long long nwall = max_face(nx, ny, nz);
            for (long long q = 0; q < n; q++) grid[q] = 1.0e10f;
            const float *w = preset;
            auto at = [&](int i, int j, int k) -> float & { return grid[(long long)k * nxy + (long long)j * nx + i]; };
            if (w[0] > -1.e-20) { long long ii = 0; for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) at(0, j, k) = w[ii++]; }
            w += nwall;
            if (w[0] > -1.e-20) { long long ii = 0; for (int k = 0; k < nz; k++) for (int j = 0; j < ny; j++) at(nx - 1, j, k) = w[ii++]; }
            w += nwall;
            if (w[0] > -1.e-20) { long long ii = 0; for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) at(i, 0, k) = w[ii++]; }
            w += nwall;
            if (w[0] > -1.e-20) { long long ii = 0; for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) at(i, ny - 1, k) = w[ii++]; }
            w += nwall;
            if (w[0] > -1.e-20) { long long ii = 0; for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) at(i, j, 0) = w[ii++]; }
            w += nwall;
            if (w[0] > -1.e-20) { long long ii = 0; for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) at(i, j, nz - 1) = w[ii++]; }
        } else {
            memcpy(grid, preset, sizeof(float) * (size_t)n);
        }
        o.init_mode = 1;
        o.radius0 = 1;
        const int sw = p.srcwall;
        if (sw == 1) x2 = 1; else { x2 = nx; dx2 = 0; }
        if (sw == 2) x1 = nx - 2; else { x1 = -1; dx1 = 0; }
        if (sw == 3) y2 = 1; else { y2 = ny; dy2 = 0; }
        if (sw == 4) y1 = ny - 2; else { y1 = -1; dy1 = 0; }
        if (sw == 5) z2 = 1; else { z2 = nz; dz2 = 0; }
        if (sw == 6) z1 = nz - 2; else { z1 = -1; dz1 = 0; }
    }
    o.box[0][0] = x1; o.box[0][1] = x2;
    o.box[1][0] = y1; o.box[1][1] = y2;
    o.box[2][0] = z1; o.box[2][1] = z2;
    o.dflag[0] = dx1; o.dflag[1] = dx2; o.dflag[2] = dy1; o.dflag[3] = dy2; o.dflag[4] = dz1; o.dflag[5] = dz2;
    o.grow[0] = p.kminus; o.grow[1] = p.kplus; o.grow[2] = p.iplus;
    o.grow[3] = p.iminus; o.grow[4] = p.jminus; o.grow[5] = p.jplus;
    o.reverse = p.reverse;
    // sphfd.c:977-1003
    int headpref = p.headpref;
    if (headpref > 0) {
        long long head = (long long)nx * ny * nz;
        if (nx > 5 && x2 <= head) { headpref = 2; head = x2; }
        if (nx > 5 && (nx - 1 - x1) <= head) { headpref = 1; head = nx - 1 - x1; }
        if (ny > 5 && y2 <= head) { headpref = 4; head = y2; }
        if (ny > 5 && (ny - 1 - y1) <= head) { headpref = 3; head = ny - 1 - y1; }
        if (nz > 5 && z2 <= head) { headpref = 6; head = z2; }
        if (nz > 5 && (nz - 1 - z1) <= head) { headpref = 5; head = nz - 1 - z1; }
    }
    o.headpref = headpref;
    return kOk;
}

}  // namespace fdprep
