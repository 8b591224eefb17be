// sweep_kernel.hip -- gfx950 kernels of the sphfd path.
//
//   sphfd_sweep_kernel   persistent: each workgroup pulls whole tables from a
//                        queue and runs the complete sweep of each
//                        (sweep_solver.h); built wide (512 lanes, one
//                        workgroup per CU) and narrow (256 lanes, two per CU)
//   sphfd_serve_kernel   persistent request loop for serving: workgroups claim
//                        requests published by the host while the kernel runs
//                        (serve_shared.h), each solved exactly like a batched
//                        table
//   slowness_kernel      1/v in FP32 (sphfd.c:689-690), model upload
//   sphfd_one_kernel     one table, its descriptor passed by value, for the
//                        entry point that allocates nothing (its device state
//                        lives in the file-scope `scratch` array below)
//
// Compiled with -ffp-contract=off: FP64 stencil arithmetic must round
// operation by operation like the reference's x86-64 build.
#include <hip/hip_runtime.h>

#include "coarse_core.h"
#include "fd_kernels.h"
#include "serve_shared.h"
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
    __device__ __forceinline__ unsigned atomic_cas(unsigned *p, unsigned cmp, unsigned v) const {
        return atomicCAS(p, cmp, v);
    }
    __device__ __forceinline__ unsigned long long atomic_max_u64(unsigned long long *p, unsigned long long v) const {
        return atomicMax(p, v);
    }
    __device__ __forceinline__ long long clock() const { return (long long)wall_clock64(); }
    // team members on other CUs (and other XCDs): agent scope
    __device__ __forceinline__ void fence_release() const { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
    __device__ __forceinline__ void fence_acquire() const { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
    __device__ __forceinline__ int load_relaxed(const int *p) const {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void store_relaxed(int *p, int v) const {
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // this wave's stores have reached the L2; also a compiler barrier, so no
    // store is scheduled after the wait
    __device__ __forceinline__ void wait_stores() const { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
    __device__ __forceinline__ void store_release(int *p, int v) const {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void pause() const { __builtin_amdgcn_s_sleep(2); }
};

// Tickets from the queue: table i < nteam has tickets team_first[i] ..
// team_first[i + 1] - 1 (its team: the first is the leader), every later
// table one ticket.  A member waits only for members that have taken their
// tickets already, so a team forms whenever its workgroups can run at once.
//
// TEAMS false (the narrow build, used when tables outnumber the CUs and teams
// never form): the team code is not compiled in.
template <int THREADS, int ARENA, int GEOL, bool TEAMS>
__global__ void __launch_bounds__(THREADS, 2)
    sphfd_sweep_kernel(const fdsweep::TableJob *jobs, int njobs, const fdsweep::Scratch *slots, int *queue,
                       fdsweep::TeamCtl *teams, const int *team_first, int nteam) {
    __shared__ fdsweep::SharedT<ARENA, GEOL> sh;
    DevExec x;
    const fdsweep::Scratch W = slots[blockIdx.x];
    if (!TEAMS) {
        while (true) {
            if (threadIdx.x == 0) sh.job = atomicAdd(queue, 1);
            __syncthreads();
            const int j = sh.job;
            __syncthreads();
            if (j >= njobs) break;
            fdsweep::solve_table(x, sh, jobs[j], W);
        }
        return;
    }
    const int tt = nteam > 0 ? team_first[nteam] : 0;
    while (true) {
        if (threadIdx.x == 0) sh.job = atomicAdd(queue, 1);
        __syncthreads();
        const int t = sh.job;
        __syncthreads();
        int j, member = 0, K = 1;
        if (t < tt) {
            int lo = 0, hi = nteam - 1;   // the last team whose first ticket is <= t
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (team_first[mid] <= t) lo = mid; else hi = mid - 1;
            }
            j = lo;
            member = t - team_first[lo];
            K = team_first[lo + 1] - team_first[lo];
        } else {
            j = nteam + (t - tt);
        }
        if (j >= njobs) break;
        if constexpr (TEAMS) {
            if (K > 1 && member > 0) {
                fdsweep::team_helper(x, sh, jobs[j], teams + j, K, member);
            } else if (K > 1) {
                fdsweep::solve_table(x, sh, jobs[j], W, teams + j, K);
                fdsweep::team_release(x, teams + j);
            } else {
                fdsweep::solve_table(x, sh, jobs[j], W);
            }
        }
    }
}

// Lane 0 of an idle workgroup: claim the next published request, or return -1
// when the host asked to stop or no workgroup has had work for `idle_ticks`
// (wall-clock ticks).  Every path of the loop re-reads the host's tail and
// stop words, so the loop ends as soon as either allows it.
__device__ long long serve_claim(fdsweep::ServeCtl *ctl, fdsweep::ServeDev *dev, long long idle_ticks) {
    while (true) {
        const unsigned long long tail = __hip_atomic_load(&ctl->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        unsigned long long h = __hip_atomic_load(&dev->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (h < tail) {
            if (__hip_atomic_compare_exchange_strong(&dev->head, &h, h + 1, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_fetch_add(&dev->busy, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_max(&dev->last, (long long)wall_clock64(), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                return (long long)h;
            }
        }
        if (__hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) return -1;
        const int busy = __hip_atomic_load(&dev->busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long long last = __hip_atomic_load(&dev->last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (busy == 0 && (long long)wall_clock64() - last > idle_ticks) return -1;
        for (int s = 0; s < 4; s++) __builtin_amdgcn_s_sleep(127);
    }
}

__global__ void __launch_bounds__(kWideThreads, 2)
    sphfd_serve_kernel(fdsweep::ServeCtl *ctl, const fdsweep::ServeJob *ring, unsigned long long *done, int cap,
                       const fdsweep::Scratch *slots, fdsweep::ServeDev *dev, long long idle_ticks) {
    __shared__ fdsweep::Shared sh;
    __shared__ fdsweep::ServeJob sj;
    __shared__ long long sjob;
    DevExec x;
    const int tid = (int)threadIdx.x, T = (int)blockDim.x;
    const fdsweep::Scratch W = slots[blockIdx.x];
    if (tid == 0)
        __hip_atomic_fetch_max(&dev->last, (long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if (tid == 0) sjob = serve_claim(ctl, dev, idle_ticks);
        __syncthreads();
        const long long j = sjob;
        if (j < 0) break;
        // the descriptor lives in host memory: one copy into LDS
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
        {
            const unsigned *src = (const unsigned *)&ring[j % cap];
            unsigned *dst = (unsigned *)&sj;
            for (int q = tid; q < (int)(sizeof(fdsweep::ServeJob) / 4); q += T) dst[q] = src[q];
        }
        __syncthreads();
        // stage the request's host inputs into its arena block
        for (int q = tid; q < sj.nsinq; q += T) sj.sinq_dst[q] = sj.sinq_src[q];
        if (sj.vel_src)
            for (long long q = tid; q < sj.nvel; q += T) sj.slow_dst[q] = 1.0f / sj.vel_src[q];
        if (sj.preset_src) {
            const long long n = (long long)sj.job.nx * sj.job.ny * sj.job.nz;
            for (long long q = tid; q < n; q += T) sj.job.time[q] = sj.preset_src[q];
        }
        if (sj.cvel_src) {
            // the request's slowness from its coarse model: coarse slowness and
            // node coordinates, per-axis cells, then every fine node (as
            // c2f_prep_kernel and c2f_fine_kernel compute them)
            const C2fGrid &cg = sj.cgrid;
            const int ncg = cg.nxc + cg.nyc + cg.nzc;
            const long long nxyzc = (long long)cg.nxc * cg.nyc * cg.nzc;
            float *gdst = (float *)cg.gx;
            for (int q = tid; q < ncg; q += T) gdst[q] = sj.cg_src[q];
            for (long long q = tid; q < nxyzc; q += T) sj.cslow_dst[q] = c2fdev::slowness(sj.cvel_src[q]);
            __syncthreads();
            const long long nax = (long long)cg.nx + cg.ny + cg.nz;
            for (long long q = tid; q < nax; q += T) sj.caxis[q] = c2fdev::axis_entry(cg, q);
            __syncthreads();
            const unsigned nx = (unsigned)cg.nx, ny = (unsigned)cg.ny, n = nx * ny * (unsigned)cg.nz;
            for (unsigned q = tid; q < n; q += T) {
                const unsigned r = q / nx, i = q - r * nx;
                const unsigned k = r / ny, j = r - k * ny;
                const float tv = c2fdev::trilinear(cg, sj.cslow_dst, sj.caxis[i], sj.caxis[nx + j], sj.caxis[nx + ny + k]);
                const float vel = 1.f / tv;             // c2f.c:570
                sj.slow_dst[q] = 1.0f / vel;           // sphfd.c:689-690
            }
        }
        __syncthreads();
        fdsweep::solve_table(x, sh, sj.job, W);
        // results out to HBM before the completion word reaches the host
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            __hip_atomic_store(&done[j % cap], (unsigned long long)j + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_fetch_max(&dev->last, (long long)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&dev->busy, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

// Geometry tables of one grid (fdsweep::geo_fill_item per row, layer and cell).
__global__ void geo_kernel(fdsweep::GeoParams p, double *out, long long n) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        fdsweep::geo_fill_item(p, e, out);
}

// Before a batched launch: the ticket counter and the control words of the
// teams' blocks (a kernel rather than a memset node: the launch is replayed
// from a captured graph).
__global__ void reset_kernel(int *q, fdsweep::TeamCtl *teams, int nteam) {
    if (threadIdx.x == 0) q[0] = 0;
    for (int t = threadIdx.x; t < nteam; t += blockDim.x) {
        fdsweep::TeamCtl &c = teams[t];
        c.seq = 0; c.release = 0; c.abort = 0; c.bar_count = 0; c.bar_gen = 0; c.inpass = 0;
        c.wait_ticks = fdsweep::kTeamWaitTicks;
    }
}

// slow may alias vel (an in-place update): each lane reads its element before
// writing it.
__global__ void slowness_kernel(const float *vel, float *slow, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        slow[i] = 1.0f / vel[i];
}

// One table solved by one wide workgroup, with the descriptor and the
// scratch layout passed by value: the kernel of fdgpu_solve_noalloc.
__global__ void __launch_bounds__(kWideThreads, 2) sphfd_one_kernel(fdsweep::TableJob J, fdsweep::Scratch W) {
    __shared__ fdsweep::Shared sh;
    DevExec x;
    fdsweep::solve_table(x, sh, J, W);
}

}  // namespace

// Device memory of fdgpu_solve_noalloc: the table's host-prepared arrays,
// counters, geometry tables and per-face scratch, carved out by the host for
// each call.  Every GPU the library runs on holds its own copy until the
// process ends; calls on one GPU share it, so calls that overlap on one GPU
// race.  Host code never uses this name as an address (on the host it is a
// stand-in symbol): fd_noalloc_scratch returns the device address.
static __device__ __align__(256) unsigned char scratch[FD_NOALLOC_SCRATCH_BYTES];

extern "C" hipError_t fd_noalloc_scratch(void **ptr) {
    // the current device's copy of the array
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

extern "C" hipError_t fd_launch_one(const fdsweep::TableJob *J, const fdsweep::Scratch *W, hipStream_t stream) {
    hipLaunchKernelGGL(sphfd_one_kernel, dim3(1), dim3(kWideThreads), 0, stream, *J, *W);
    return hipGetLastError();
}

extern "C" hipError_t fd_launch_sweep(int variant, const fdsweep::TableJob *d_jobs, int njobs,
                                      const fdsweep::Scratch *d_slots, int nslots, int *d_queue,
                                      fdsweep::TeamCtl *d_teams, const int *d_team_first, int nteam,
                                      hipStream_t stream) {
    if (nteam <= 0 || !d_teams || !d_team_first) nteam = 0;
    if (variant == FD_NARROW)
        hipLaunchKernelGGL((sphfd_sweep_kernel<kNarrowThreads, kNarrowArenaWords, kNarrowGeoLds, false>), dim3(nslots),
                           dim3(kNarrowThreads), 0, stream, d_jobs, njobs, d_slots, d_queue, d_teams, d_team_first, 0);
    else
        hipLaunchKernelGGL((sphfd_sweep_kernel<kWideThreads, fdsweep::kArenaWords, fdsweep::kGeoLdsDoubles, true>), dim3(nslots), dim3(kWideThreads), 0,
                           stream, d_jobs, njobs, d_slots, d_queue, d_teams, d_team_first, nteam);
    return hipGetLastError();
}

extern "C" hipError_t fd_launch_geo(const fdsweep::GeoParams *p, double *d_out, hipStream_t stream) {
    const long long n = fdsweep::geo_items(p->ny, p->nz);
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(geo_kernel, dim3(blocks), dim3(256), 0, stream, *p, d_out, n);
    return hipGetLastError();
}

extern "C" hipError_t fd_launch_serve(fdsweep::ServeCtl *ctl, const fdsweep::ServeJob *ring, unsigned long long *done,
                                      int cap, const fdsweep::Scratch *d_slots, int nslots, fdsweep::ServeDev *dev,
                                      long long idle_ticks, hipStream_t stream) {
    hipLaunchKernelGGL(sphfd_serve_kernel, dim3(nslots), dim3(kWideThreads), 0, stream, ctl, ring, done, cap,
                       d_slots, dev, idle_ticks);
    return hipGetLastError();
}

extern "C" const void *fd_sweep_kernel_ptr(int variant) {
    return variant == FD_NARROW
               ? (const void *)sphfd_sweep_kernel<kNarrowThreads, kNarrowArenaWords, kNarrowGeoLds, false>
               : (const void *)sphfd_sweep_kernel<kWideThreads, fdsweep::kArenaWords, fdsweep::kGeoLdsDoubles, true>;
}

extern "C" hipError_t fd_launch_reset(int *d_queue, fdsweep::TeamCtl *d_teams, int nteam, hipStream_t stream) {
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, d_queue, d_teams, nteam);
    return hipGetLastError();
}

extern "C" hipError_t fd_launch_slowness(const float *d_vel, float *d_slow, long long n, hipStream_t stream) {
    const int threads = 256;
    long long blocks = (n + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(slowness_kernel, dim3((unsigned)blocks), dim3(threads), 0, stream, d_vel, d_slow, n);
    return hipGetLastError();
}
