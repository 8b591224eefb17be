// Host build of the device ray tracer (fdtomoc_amd/csrc/ray_core.h) for the
// CPU tests: the same per-lane code with the host's math library and a dense
// derivative array, exported with the oracle's signature
// (oracle/ray_oracle.h, oracle_ray_trace_many) so the two can be compared
// ray for ray.  Test infrastructure only.
#include <string.h>

#include <vector>

#include "../../fdtomoc_amd/csrc/ray_prep.h"
#include "../../oracle/ray_oracle.h"

namespace {

int g_math = 0;   // 0: the host C library; 1: the device's correctly rounded functions

struct DenseAcc {
    float *du;
    int nvar;
    void add(int v, double c) {
        if (v >= 0 && v < nvar) du[v] = (float)((double)du[v] + c);
    }
};

}  // namespace

extern "C" void emu_ray_set_math(int m) { g_math = m; }

extern "C" double emu_cr(int fn, double a, double b) {
    double s, c;
    switch (fn) {
    case 0: crm::cr_sincos(a, s, c); return s;
    case 1: crm::cr_sincos(a, s, c); return c;
    default: return crm::cr_atan2(a, b);
    }
}

extern "C" int emu_ray_trace_many(const oracle_ray_grid *g, const float *const *tables, const oracle_ray_in *rays,
                                  int n, oracle_ray_out *out, int *idx, float *val, int cap, int nthreads) {
    (void)nthreads;
    std::vector<double> trig = fdray::face_trig(g->nx, g->ny, g->x0, g->y0, g->df, g->dq);
    const std::vector<int> mx = fdray::coarse_guess(g->gx, g->nxc, g->x0, g->df, g->nx),
                           my = fdray::coarse_guess(g->gy, g->nyc, g->y0, g->dq, g->ny),
                           mz = fdray::coarse_guess(g->gz, g->nzc, g->z0, g->h, g->nz);
    fdray::RayGrid G{g->nx, g->ny, g->nz, g->nxc, g->nyc, g->nzc, g->ido1d, g->max_segments,
                     g->x0, g->y0, g->z0, g->df, g->dq, g->h, g->gx, g->gy, g->gz, trig.data(),
                     mx.data(), my.data(), mz.data()};
    const int nvar = g->ido1d ? g->nzc : g->nxc * g->nyc * g->nzc;
    std::vector<float> du((size_t)nvar);
    int over = 0;
    for (int r = 0; r < n; r++) {
        fdray::RayIn in{rays[r].ex, rays[r].ey, rays[r].ez, rays[r].xs, rays[r].ys, rays[r].zs, 0, 0, 0, 0};
        fdray::station_point(in.xs, in.ys, in.zs, in.x2, in.y2, in.z2);
        std::fill(du.begin(), du.end(), 0.0f);
        DenseAcc acc{du.data(), nvar};
        fdray::RayOut o;
        if (g_math == 1)
            fdray::ray_trace<fdray::CrMath>(G, tables[r], in, o, acc);
        else
            fdray::ray_trace<fdray::LibMath>(G, tables[r], in, o, acc);
        oracle_ray_out &q = out[r];
        q.status = o.status;
        q.nseg = o.nseg;
        q.first_in_cube = o.first_in_cube;
        q.tev = o.tev;
        q.dtdx = o.dtdx; q.dtdy = o.dtdy; q.dtdz = o.dtdz;
        fdray::take_off(o, q.ai, q.az);
        q.rdev = o.rdev;
        int k = 0;
        if (o.status == fdray::kRayOk)
            for (int v = 0; v < nvar; v++)
                if (du[(size_t)v] != 0) {
                    if (k < cap) {
                        idx[(size_t)r * cap + k] = v;
                        val[(size_t)r * cap + k] = du[(size_t)v];
                    }
                    k++;
                }
        q.ndu = k;
        if (k > cap) over++;
    }
    return over;
}
