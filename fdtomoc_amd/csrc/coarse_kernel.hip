// coarse_kernel.hip -- gfx950 kernels of the coarse-to-fine model
// interpolation of FDtomoC's c2f (src/c2f.c:549-575, find_vel 673-727): the
// fine velocity grids that sphfd consumes, computed on the device from an
// inversion's coarse model and written straight into resident models.
//
//   c2f_prep_kernel  coarse velocities -> slowness (a double quotient rounded
//                    to float, c2f.c:549-552); per fine coordinate of each
//                    axis the coarse cell and the fraction across it
//   c2f_fine_kernel  fine nodes across the lanes, P and S: trilinear slowness in
//                    the reference's float operation order, velocity 1/v,
//                    the model's slowness 1/velocity (sphfd.c:689-690), and
//                    each workgroup's partial velocity range per phase
//   c2f_range_kernel one workgroup: the partial ranges folded
//
// Compiled with -ffp-contract=off: every float multiply and add rounds
// separately, as in the reference's x86-64 build.
#include <hip/hip_runtime.h>

#include "coarse_core.h"

namespace {

__global__ void c2f_prep_kernel(const C2fGrid grid, float *coarse, long long ncoarse, C2fAxis *axis) {
    const long long t0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = t0; q < ncoarse; q += stride) coarse[q] = c2fdev::slowness(coarse[q]);
    const long long nax = (long long)grid.nx + grid.ny + grid.nz;
    for (long long q = t0; q < nax; q += stride) axis[q] = c2fdev::axis_entry(grid, q);
}

// Block-wide min/max of positive finite velocities (as ordered bit patterns)
// and a flag for any other value, stored as this block's partial range
// part[0..2] (plain stores: c2f_range_kernel folds the blocks).
__device__ void fold_range(unsigned lo, unsigned hi, unsigned bad, unsigned *part) {
    __shared__ unsigned slo[64], shi[64], sbad[64];
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, off));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, off));
        bad |= (unsigned)__shfl_xor((int)bad, off);
    }
    const int w = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
    if ((threadIdx.x & 63) == 0) { slo[w] = lo; shi[w] = hi; sbad[w] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < nw; q++) { lo = min(lo, slo[q]); hi = max(hi, shi[q]); bad |= sbad[q]; }
        part[0] = lo;
        part[1] = hi;
        part[2] = bad;
    }
    __syncthreads();
}

// One block: the fine kernel's nblk partial ranges (8 words each: P at 0..2,
// S at 4..6) folded into range[0..2] and range[4..6].
__global__ void c2f_range_kernel(const unsigned *__restrict__ part, int nblk, unsigned *range) {
    __shared__ unsigned s[6][4];
    unsigned v[6] = {0xffffffffu, 0u, 0u, 0xffffffffu, 0u, 0u};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const unsigned *p = part + 8 * (size_t)b;
        v[0] = min(v[0], p[0]); v[1] = max(v[1], p[1]); v[2] |= p[2];
        v[3] = min(v[3], p[4]); v[4] = max(v[4], p[5]); v[5] |= p[6];
    }
    for (int off = 32; off > 0; off >>= 1) {
        v[0] = min(v[0], (unsigned)__shfl_xor((int)v[0], off));
        v[1] = max(v[1], (unsigned)__shfl_xor((int)v[1], off));
        v[2] |= (unsigned)__shfl_xor((int)v[2], off);
        v[3] = min(v[3], (unsigned)__shfl_xor((int)v[3], off));
        v[4] = max(v[4], (unsigned)__shfl_xor((int)v[4], off));
        v[5] |= (unsigned)__shfl_xor((int)v[5], off);
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 6; q++) s[q][w] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < (int)(blockDim.x + 63) / 64; x++) {
            v[0] = min(v[0], s[0][x]); v[1] = max(v[1], s[1][x]); v[2] |= s[2][x];
            v[3] = min(v[3], s[3][x]); v[4] = max(v[4], s[4][x]); v[5] |= s[5][x];
        }
        range[0] = v[0]; range[1] = v[1]; range[2] = v[2];
        range[4] = v[3]; range[5] = v[4]; range[6] = v[5];
    }
}

__global__ void c2f_fine_kernel(const C2fGrid grid, const float *__restrict__ sl, const C2fAxis *__restrict__ axis,
                                float *__restrict__ vel_out, float *__restrict__ slow_p, float *__restrict__ slow_s,
                                unsigned *part) {
    // fine nodes in file order (x fastest) across the lanes: coalesced writes;
    // the host keeps nx*ny*nz below 2^31
    const unsigned nx = (unsigned)grid.nx, ny = (unsigned)grid.ny;
    const unsigned n = nx * ny * (unsigned)grid.nz;
    const unsigned nxyzc = (unsigned)(grid.nxc * grid.nyc * grid.nzc);
    unsigned lo[2] = {0xffffffffu, 0xffffffffu}, hi[2] = {0u, 0u}, bad[2] = {0u, 0u};
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const unsigned r = q / nx, i = q - r * nx;
        const unsigned k = r / ny, j = r - k * ny;
        const C2fAxis ex = axis[i], ey = axis[nx + j], ez = axis[nx + ny + k];
#pragma unroll
        for (int ph = 0; ph < 2; ph++) {
            const float tv = c2fdev::trilinear(grid, sl + ph * nxyzc, ex, ey, ez);
            const float vel = 1.f / tv;                     // c2f.c:570
            if (vel_out) vel_out[(size_t)ph * n + q] = vel;
            (ph == 0 ? slow_p : slow_s)[q] = 1.0f / vel;    // sphfd.c:689-690
            unsigned u;
            __builtin_memcpy(&u, &vel, 4);
            if (vel > 0.f && vel < __builtin_inff()) {
                lo[ph] = min(lo[ph], u);
                hi[ph] = max(hi[ph], u);
            } else {
                bad[ph] = 1u;
            }
        }
    }
    fold_range(lo[0], hi[0], bad[0], part + 8 * (size_t)blockIdx.x);
    fold_range(lo[1], hi[1], bad[1], part + 8 * (size_t)blockIdx.x + 4);
}

}  // namespace

extern "C" hipError_t fd_launch_c2f(const C2fGrid *grid, float *d_coarse, C2fAxis *d_axis, float *d_vel_out,
                                    float *d_slow_p, float *d_slow_s, unsigned *d_part, unsigned *d_range,
                                    hipStream_t stream) {
    const long long ncoarse = 2LL * grid->nxc * grid->nyc * grid->nzc;
    const long long nax = (long long)grid->nx + grid->ny + grid->nz;
    const long long m = ncoarse > nax ? ncoarse : nax;
    int blocks = (int)((m + 255) / 256);
    hipLaunchKernelGGL(c2f_prep_kernel, dim3(blocks), dim3(256), 0, stream, *grid, d_coarse, ncoarse, d_axis);
    // kC2fBlocksPerCu blocks per CU striding over the nodes (each stores its
    // partial range; c2f_range_kernel folds them).  MI355X, 301x301x61 per
    // call: 1 per CU 0.075 ms, 2 0.055, 4 0.051, 8 0.052, 16 0.055 (both
    // kernels); with the range folded by atomics instead, 2 per CU was best
    // at 0.067 ms and 2048 blocks took 0.093
    const long long n = (long long)grid->nx * grid->ny * grid->nz;
    if (n >= (1LL << 31)) return hipErrorInvalidValue;
    const long long nb = (n + 255) / 256;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    long long cap = (long long)kC2fBlocksPerCu * cus;
    if (cap > kC2fMaxBlocks) cap = kC2fMaxBlocks;
    const int nblk = (int)(nb < cap ? nb : cap);
    hipLaunchKernelGGL(c2f_fine_kernel, dim3((unsigned)nblk), dim3(256), 0, stream, *grid, (const float *)d_coarse,
                       (const C2fAxis *)d_axis, d_vel_out, d_slow_p, d_slow_s, d_part);
    hipLaunchKernelGGL(c2f_range_kernel, dim3(1), dim3(256), 0, stream, (const unsigned *)d_part, nblk, d_range);
    return hipGetLastError();
}
