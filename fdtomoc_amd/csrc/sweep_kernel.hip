// sweep_kernel.hip -- gfx950 kernels of the sphfd path.
//
//   sphfd_sweep_kernel   persistent: each workgroup pulls whole tables from a
//                        queue and runs the complete sweep of each
//                        (sweep_solver.h) with 256 lanes
//   slowness_kernel      1/v in FP32 (sphfd.c:689-690), model upload
//
// Compiled with -ffp-contract=off: FP64 stencil arithmetic must round
// operation by operation like the reference's x86-64 build.
#include <hip/hip_runtime.h>

#include "fd_kernels.h"
#include "sweep_solver.h"

namespace {

struct DevExec {
    __device__ __forceinline__ int tid() const { return (int)threadIdx.x; }
    __device__ __forceinline__ int nthr() const { return (int)blockDim.x; }
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ int atomic_add(int *p, int v) const { return atomicAdd(p, v); }
    __device__ __forceinline__ int atomic_exch(int *p, int v) const { return atomicExch(p, v); }
    __device__ __forceinline__ unsigned atomic_sub_u(unsigned *p, unsigned v) const { return atomicSub(p, v); }
    __device__ __forceinline__ unsigned atomic_or(unsigned *p, unsigned v) const { return atomicOr(p, v); }
    __device__ __forceinline__ unsigned atomic_and(unsigned *p, unsigned v) const { return atomicAnd(p, v); }
    __device__ __forceinline__ long long clock() const { return (long long)wall_clock64(); }
};

__global__ void __launch_bounds__(FD_SWEEP_THREADS)
    sphfd_sweep_kernel(const fdsweep::TableJob *jobs, int njobs, const fdsweep::Scratch *slots, int *queue) {
    __shared__ fdsweep::Shared sh;
    DevExec x;
    const fdsweep::Scratch W = slots[blockIdx.x];
    while (true) {
        if (threadIdx.x == 0) sh.job = atomicAdd(queue, 1);
        __syncthreads();
        const int j = sh.job;
        __syncthreads();
        if (j >= njobs) break;
        fdsweep::solve_table(x, sh, jobs[j], W);
    }
}

__global__ void slowness_kernel(const float *__restrict__ vel, float *__restrict__ slow, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        slow[i] = 1.0f / vel[i];
}

}  // namespace

extern "C" hipError_t fd_launch_sweep(const fdsweep::TableJob *d_jobs, int njobs, const fdsweep::Scratch *d_slots,
                                      int nslots, int *d_queue, hipStream_t stream) {
    hipLaunchKernelGGL(sphfd_sweep_kernel, dim3(nslots), dim3(FD_SWEEP_THREADS), 0, stream, d_jobs, njobs, d_slots,
                       d_queue);
    return hipGetLastError();
}

extern "C" const void *fd_sweep_kernel_ptr() { return (const void *)sphfd_sweep_kernel; }

extern "C" hipError_t fd_launch_slowness(const float *d_vel, float *d_slow, long long n, hipStream_t stream) {
    const int threads = 256;
    long long blocks = (n + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(slowness_kernel, dim3((unsigned)blocks), dim3(threads), 0, stream, d_vel, d_slow, n);
    return hipGetLastError();
}
