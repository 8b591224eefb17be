// ray_kernels.h -- launch wrapper of the ray-tracing kernel of FDtomoC's
// sphrayderv (ray_core.h) over resident travel-time tables, shared by the
// kernel (ray_kernel.hip) and the host side (ray.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "ray_core.h"

// Device view of one ray's result: RayOut plus where its derivatives went in
// the launch's packed arrays (-1: they did not fit; the host traces the ray
// again in a later launch).
struct RayResult {
    fdray::RayOut o;
    long long offset;
};

// Traces rays [0, n): lane r traces ray r through tables[rays[r].table],
// accumulating its derivatives in row r of du (rows of dstride >= nvar
// floats) with a bit per touched variable in row r of bits (rows of bstride
// >= (nvar + 31) / 32 words), both all zero on entry and on return, then
// appends its nonzero derivatives in index order to (idx, val) at an offset
// taken from *cursor, up to `capacity` entries in all.
extern "C" hipError_t fd_launch_rays(const fdray::RayGrid *grid, const fdray::RayIn *rays, const float *const *tables,
                                     int n, float *du, unsigned *bits, int nvar, long long dstride, long long bstride,
                                     RayResult *results,
                                     unsigned long long *cursor, int *idx, float *val, long long capacity,
                                     hipStream_t stream);

// Copies each ray's packed derivatives to its place in output order: ray k of
// the launch (results[k]) has results[k].o.ndu entries at results[k].offset
// of (idx, val), which go to dst[k] of (out_idx, out_val).  One wave per ray.
extern "C" hipError_t fd_launch_ray_place(const RayResult *results, const long long *dst, int n, const int *idx,
                                          const float *val, int *out_idx, float *out_val, hipStream_t stream);
