// coarse_core.h -- device functions of c2f's coarse-to-fine interpolation
// (src/c2f.c:549-575, find_vel 673-727), shared by the batched kernels
// (coarse_kernel.hip) and by the serving loop, which interpolates a
// request's own slowness from the coarse model it carries
// (sphfd_serve_kernel in sweep_kernel.hip).  Float arithmetic in the
// reference's operation order; compiled without FP contraction.
#pragma once

#include <hip/hip_runtime.h>

#include "coarse_kernels.h"

namespace c2fdev {

// The coarse cell holding v on axis g (c2f.c:675-683): the first node
// strictly beyond v, clamped to the last interval, minus one.
__device__ __forceinline__ int cell(const float *g, int n, float v) {
    int i;
    for (i = 1; i < n; i++)
        if (g[i] > v) break;
    if (i == n) i--;
    return i - 1;
}

// Coarse slowness of a velocity: a double quotient rounded to float (c2f.c:549-552).
__device__ __forceinline__ float slowness(float v) { return (float)(1. / (double)v); }

// Entry q of the per-axis table: axes x (q < nx), y, z in turn, each fine
// coordinate o + i h rounded to float (c2f.c:560-566), its cell and its
// fraction across the cell (c2f.c:698-704).
__device__ __forceinline__ C2fAxis axis_entry(const C2fGrid &grid, long long q) {
    const int a = q < grid.nx ? 0 : q < grid.nx + grid.ny ? 1 : 2;
    const int i = (int)(a == 0 ? q : a == 1 ? q - grid.nx : q - grid.nx - grid.ny);
    const double o = a == 0 ? grid.x0 : a == 1 ? grid.y0 : grid.z0;
    const float *g = a == 0 ? grid.gx : a == 1 ? grid.gy : grid.gz;
    const int nc = a == 0 ? grid.nxc : a == 1 ? grid.nyc : grid.nzc;
    const float v = (float)(o + i * grid.h);
    C2fAxis e;
    e.cell = cell(g, nc, v);
    const float gi = g[e.cell];
    const float hg = g[e.cell + 1] - gi;
    e.f = (v - gi) / hg;
    return e;
}

// Trilinear slowness at fine node (i, j, k) from one phase's coarse
// slowness s (find_vel's sum, c2f.c:711-724, term by term).
__device__ __forceinline__ float trilinear(const C2fGrid &grid, const float *__restrict__ s, C2fAxis ex, C2fAxis ey,
                                           C2fAxis ez) {
    const int ci = ex.cell, cj = ey.cell, ck = ez.cell;
    const int nk = grid.nxc * grid.nyc * ck;
    const int nj = grid.nxc * cj;
    const int nk2 = grid.nxc * grid.nyc * (ck + 1);
    const int nj2 = grid.nxc * (cj + 1);
    const float fx = ex.f, fy = ey.f, fz = ez.f;
    s += ci;
    float tv = 0.f;
    tv = (1.f - fx) * (1.f - fy) * (1.f - fz) * s[nk + nj];
    tv = tv + fx * (1.f - fy) * (1.f - fz) * s[nk + nj + 1];
    tv = tv + (1.f - fx) * fy * (1.f - fz) * s[nk + nj2];
    tv = tv + (1.f - fx) * (1.f - fy) * fz * s[nk2 + nj];
    tv = tv + fx * fy * (1.f - fz) * s[nk + nj2 + 1];
    tv = tv + fx * (1.f - fy) * fz * s[nk2 + nj + 1];
    tv = tv + (1.f - fx) * fy * fz * s[nk2 + nj2];
    tv = tv + fx * fy * fz * s[nk2 + nj2 + 1];
    return tv;
}

}  // namespace c2fdev
