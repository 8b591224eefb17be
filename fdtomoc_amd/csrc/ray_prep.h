// ray_prep.h -- host-side preparation of a ray launch (ray_core.h): the
// per-row and per-column sines and cosines of the grid's cell faces and each
// station's Cartesian point, computed with the host C library so that they
// are the values the reference program computes (sphrayderv.c:1790-1795,
// 1833-1836, 1895-1945).
#pragma once

#include <math.h>

#include <vector>

#include "ray_core.h"

namespace fdray {

// Layout of RayGrid::trig: rows j = -1..ny, then columns i = -1..nx, four
// doubles each.
inline std::vector<double> face_trig(int nx, int ny, double x0, double y0, double df, double dq) {
    std::vector<double> v(4 * (size_t)(ny + 2) + 4 * (size_t)(nx + 2));
    size_t o = 0;
    for (int j = -1; j <= ny; j++) {
        const double yj = dq * j + y0;   // sphrayderv.c:1470
        v[o++] = sin(yj);
        v[o++] = cos(yj);
        v[o++] = sin(yj + dq);
        v[o++] = cos(yj + dq);
    }
    for (int i = -1; i <= nx; i++) {
        const double xi = df * i + x0;   // sphrayderv.c:1469
        v[o++] = sin(xi);
        v[o++] = cos(xi);
        v[o++] = sin(xi + df);
        v[o++] = cos(xi + df);
    }
    return v;
}

// The station's point as the reference converts it for the straight rays
// of the station cube (sphrayderv.c:1832-1836).
inline void station_point(double xs, double ys, double zs, float &x2, float &y2, float &z2) {
    const double r2 = kREarth - zs;
    x2 = (float)(r2 * sin(ys) * cos(xs));
    y2 = (float)(r2 * sin(ys) * sin(xs));
    z2 = (float)(r2 * cos(ys));
}

// Guesses for the coarse searches (RayGrid::mx/my/mz): per fine node f of
// an axis, the search's result at the middle of the fine cell f.
inline std::vector<int> coarse_guess(const float *g, int n, double g0, double step, int nf) {
    std::vector<int> m((size_t)nf);
    for (int f = 0; f < nf; f++) {
        const float v = (float)(g0 + (f + 0.5) * step);
        int i = 1;
        while (i < n && !(v < g[i])) i++;
        m[(size_t)f] = i;
    }
    return m;
}

// The take-off angles of a ray (sphrayderv.c:1865-1873): azimuth and
// incidence in degrees from its first step's direction, with the host's
// atan2f as the reference computes them; NaN when the first step ran in the
// station cube (the reference then reports its previous ray's values).
inline void take_off(const RayOut &o, float &ai, float &az) {
    if (o.first_in_cube || o.status == kRayEventOut || o.status == kRayUnreached || o.nseg == 0) {
        ai = az = __builtin_nanf("");
        return;
    }
    const double sfq = sqrt(o.sf0 * o.sf0 + o.sq0 * o.sq0);
    az = (float)(atan2f((float)o.sf0, (float)-o.sq0) / kDegRad);
    ai = (float)(atan2f((float)sfq, (float)-o.sr0) / kDegRad);
}

}  // namespace fdray
