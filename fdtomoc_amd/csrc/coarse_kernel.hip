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
//                    the velocity range of each phase
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
// and a flag for any other value, folded into range[4 * phase + 0..2].
__device__ void fold_range(unsigned lo, unsigned hi, unsigned bad, unsigned *range) {
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
        // the words only move one way: skip the atomics that cannot change them
        if (lo < __hip_atomic_load(&range[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(&range[0], lo);
        if (hi > __hip_atomic_load(&range[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&range[1], hi);
        if (bad) atomicOr(&range[2], 1u);
    }
    __syncthreads();
}

__global__ void c2f_fine_kernel(const C2fGrid grid, const float *__restrict__ sl, const C2fAxis *__restrict__ axis,
                                float *__restrict__ vel_out, float *__restrict__ slow_p, float *__restrict__ slow_s,
                                unsigned *range) {
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
    fold_range(lo[0], hi[0], bad[0], range);
    fold_range(lo[1], hi[1], bad[1], range + 4);
}

}  // namespace

extern "C" hipError_t fd_launch_c2f(const C2fGrid *grid, float *d_coarse, C2fAxis *d_axis, float *d_vel_out,
                                    float *d_slow_p, float *d_slow_s, unsigned *d_range, hipStream_t stream) {
    const long long ncoarse = 2LL * grid->nxc * grid->nyc * grid->nzc;
    const long long nax = (long long)grid->nx + grid->ny + grid->nz;
    const long long m = ncoarse > nax ? ncoarse : nax;
    int blocks = (int)((m + 255) / 256);
    hipLaunchKernelGGL(c2f_prep_kernel, dim3(blocks), dim3(256), 0, stream, *grid, d_coarse, ncoarse, d_axis);
    // two blocks per CU striding over the nodes (each folds its range once):
    // on MI355X, 384-512 blocks gave 0.067 ms per 301x301x61 update against
    // 0.093 ms with 2048 and 0.37 ms with one block per 256 nodes
    const long long n = (long long)grid->nx * grid->ny * grid->nz;
    if (n >= (1LL << 31)) return hipErrorInvalidValue;
    const long long nb = (n + 255) / 256;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    const long long cap = 2LL * cus;
    hipLaunchKernelGGL(c2f_fine_kernel, dim3((unsigned)(nb < cap ? nb : cap)), dim3(256), 0, stream, *grid,
                       (const float *)d_coarse, (const C2fAxis *)d_axis, d_vel_out, d_slow_p, d_slow_s, d_range);
    return hipGetLastError();
}
