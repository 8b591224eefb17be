// serve_shared.h -- layout shared by the serving host code (serve.cpp) and the
// device-resident request loop (sphfd_serve_kernel in sweep_kernel.hip).
//
// The host publishes requests in a ring of ServeJob descriptors kept in
// fine-grained (coherent) pinned host memory and bumps ServeCtl::tail.  Idle
// workgroups of the server kernel claim request numbers from a device counter,
// copy the descriptor into LDS, stage the request's host-side inputs into its
// HBM arena block, solve the table and post the request number + 1 into
// done[seq % capacity] (also pinned host memory).  Everything a request reads
// from device memory is written by the workgroup that solves it, or existed
// before the server kernel was launched.
#pragma once

#include "coarse_kernels.h"
#include "sweep_solver.h"

namespace fdsweep {

struct ServeCtl {
    unsigned long long tail;   // requests published by the host
    int stop;                  // 1: workgroups leave as soon as they are idle
    int pad;
};

struct ServeJob {
    TableJob job;              // device pointers into the request's arena block
    const float *vel_src;      // host: the request's own velocities (nullable)
    float *slow_dst;           // device: their slowness (job.slow points here)
    long long nvel;
    const double *sinq_src;    // host: sin(qc(j)), j = -1..ny
    double *sinq_dst;          // device copy (job.sinq points here)
    int nsinq;
    int pad;
    const float *preset_src;   // host: srctype 2/3 initial grid (nullable)
    // A request that carries one phase of a coarse model instead of
    // velocities: the workgroup interpolates its slowness into slow_dst
    // (c2f's interpolation, src/c2f.c:549-575) before solving.
    const float *cvel_src;     // host: nxc*nyc*nzc coarse velocities (nullable)
    const float *cg_src;       // host: coarse node coordinates gx | gy | gz
    float *cslow_dst;          // device: the coarse slowness
    C2fAxis *caxis;            // device: nx + ny + nz axis entries
    C2fGrid cgrid;             // gx/gy/gz point at the device copy of cg_src
};

// Device counters of the server (one allocation).
struct ServeDev {
    unsigned long long head;   // next request number to claim
    int busy;                  // workgroups inside a request
    int pad;
    long long last;            // wall clock of the last claim or completion
};

}  // namespace fdsweep
