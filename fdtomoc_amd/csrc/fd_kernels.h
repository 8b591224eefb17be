// fd_kernels.h -- launch wrappers of the gfx950 kernels (sweep_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "serve_shared.h"
#include "sweep_solver.h"

#ifndef FD_SWEEP_THREADS
#define FD_SWEEP_THREADS 512   // lanes per workgroup of the sweep and server kernels (2 waves per SIMD)
#endif

extern "C" hipError_t fd_launch_sweep(const fdsweep::TableJob *d_jobs, int njobs, const fdsweep::Scratch *d_slots,
                                      int nslots, int *d_queue, hipStream_t stream);
extern "C" hipError_t fd_launch_slowness(const float *d_vel, float *d_slow, long long n, hipStream_t stream);
extern "C" hipError_t fd_launch_serve(fdsweep::ServeCtl *ctl, const fdsweep::ServeJob *ring, unsigned long long *done,
                                      int cap, const fdsweep::Scratch *d_slots, int nslots, fdsweep::ServeDev *dev,
                                      long long idle_ticks, hipStream_t stream);
