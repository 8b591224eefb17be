// loc_geo.cpp -- the host-side geography of sphfdloc (include/fdtomoc_amd.h,
// fdgpu_loc_params_from_grid and fdgpu_loc_geographic): the search grid as
// sphfdloc sets it up from the spec (src/sphfdloc.c:382-393) and the
// hypocentre's geographic latitude, longitude and depth as it prints them
// (sphfdloc.c:1131-1133).  The WGS84 conversions follow
// lib/common/geographic_method.c (glath, glathinv); sphfdloc's constants
// are float literals widened to double (sphfdloc.c:25-26).
#include <math.h>

#include "../../include/fdtomoc_amd.h"

namespace {

const double kHpi = (double)1.570796f;
const double kDegrad = (double)0.017453292f;
const double kRearth = 6371.0;
const double kAp = 6378137.0;                       // WGS84 semi-major axis, m
const double kEsq = 6.69437978616733379 * 0.001;    // first eccentricity squared

// Geodetic latitude (rad) and ellipsoidal depth (km) -> geocentric latitude
// (rad) and radius (km) (geographic_method.c:60-83).
double glath(double xlat, double h, double *r) {
    const double s = sin(xlat);
    const double hm = -h * 1000.0;
    const double anu = kAp / sqrt(1.0 - kEsq * s * s);
    const double x = (anu + hm) * cos(xlat);
    const double z = ((1.0 - kEsq) * anu + hm) * s;
    *r = sqrt(x * x + z * z) / 1000.0;
    return atan2(z, x);
}

// Geocentric latitude (rad) and radius (km) -> geodetic latitude (rad) and
// ellipsoidal depth (km), five fixed-point steps (geographic_method.c:107-140).
double glathinv(double xcent, double r, double *h) {
    const double rm = r * 1000.0;
    double lat = xcent;
    const double x = rm * cos(lat);
    const double z = rm * sin(lat);
    double anu = 0.0;
    for (int it = 0; it < 5; it++) {
        const double s = sin(lat);
        anu = kAp / sqrt(1.0 - kEsq * s * s);
        const double hb = x / cos(lat) - anu;
        const double zp = z / (1.0 - kEsq * (anu / (anu + hb)));
        lat = atan2(zp, x);
    }
    *h = -(x / cos(lat) - anu) / 1000.0;
    return lat;
}

}  // namespace

extern "C" int fdgpu_loc_params_from_grid(int nx, int ny, int nz, double x0_deg, double y0_deg, double z0, double h,
                                          double dq_deg, double df_deg, fdgpu_loc_params *p, double *z0r) {
    if (!p || !z0r || nx < 2 || ny < 2 || nz < 2 || !(h > 0.0)) return FDGPU_EINPUT;
    fdgpu_loc_params_defaults(p);
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->z0 = z0;
    p->h = h;
    p->y0 = kHpi - glath(y0_deg * kDegrad, z0, z0r);
    p->x0 = x0_deg * kDegrad;
    p->dq = dq_deg * kDegrad;
    p->df = df_deg * kDegrad;
    if (p->dq == 0) p->dq = h / kRearth;
    if (p->df == 0) p->df = fabs(h / (kRearth * sin(p->y0)));
    return FDGPU_OK;
}

extern "C" int fdgpu_loc_geographic(double exm, double eym, double ezm, double z0, double z0r, double *lat_deg,
                                    double *lon_deg, double *depth) {
    if (!lat_deg || !lon_deg || !depth) return FDGPU_EINPUT;
    *lon_deg = exm / kDegrad;
    *lat_deg = glathinv(kHpi - eym, z0r - (ezm - z0), depth) / kDegrad;
    return FDGPU_OK;
}
