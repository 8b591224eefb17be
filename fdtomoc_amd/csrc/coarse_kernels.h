// coarse_kernels.h -- launch wrapper of the coarse-to-fine model kernels
// (coarse_kernel.hip), shared with the host side (coarse.cpp).
#pragma once

#include <hip/hip_runtime.h>

// A coarse model's node coordinates (device pointers, float as c2f.c:59
// declares them) and the fine grid they span.
struct C2fGrid {
    int nxc, nyc, nzc;
    const float *gx, *gy, *gz;
    int nx, ny, nz;
    double x0, y0, z0, h;
};

// One fine coordinate of an axis: its coarse cell and its fraction across
// the cell, (v - g[cell]) / (g[cell + 1] - g[cell]) in float (c2f.c:698-704).
struct C2fAxis {
    int cell;
    float f;
};

// Workgroups of the fine kernel: kC2fBlocksPerCu per CU, at most kC2fMaxBlocks
// (the partial-range scratch holds 8 words per workgroup).
constexpr int kC2fBlocksPerCu = 4;
constexpr int kC2fMaxBlocks = 4096;

// Converts d_coarse (2*nxc*nyc*nzc velocities, P then S) to slowness in
// place, fills d_axis (nx + ny + nz entries), then writes the fine P and S
// slownesses (and the fine velocities to d_vel_out when it is not null) and
// the velocity range into d_range[0..2] (P) and [4..6] (S): min and max bit
// patterns of positive finite velocities, and a flag for any other value.
// d_part: 8 * kC2fMaxBlocks words of scratch for the workgroups' partials.
extern "C" hipError_t fd_launch_c2f(const C2fGrid *grid, float *d_coarse, C2fAxis *d_axis, float *d_vel_out,
                                    float *d_slow_p, float *d_slow_s, unsigned *d_part, unsigned *d_range,
                                    hipStream_t stream);
