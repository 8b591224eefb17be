// fd_kernels.h -- launch wrappers of the gfx950 kernels (sweep_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "serve_shared.h"
#include "sweep_solver.h"

// Two builds of the sweep kernel, both at two wavefronts per SIMD:
//   wide    512 lanes, one workgroup per CU (LDS arena of kArenaWords words,
//           geometry tables up to kGeoLdsDoubles): the fastest per table, used
//           while there are no more tables than CUs
//   narrow  256 lanes, two workgroups per CU (LDS arena of kNarrowArenaWords,
//           geometry tables up to kNarrowGeoLds): two tables share a CU, used
//           for larger batches
// The server kernel is built wide.
enum { FD_WIDE = 0, FD_NARROW = 1 };
constexpr int kWideThreads = 512, kNarrowThreads = 256;
constexpr int kNarrowArenaWords = 13400;
constexpr int kNarrowGeoLds = 3400;   // geometry row/layer doubles in LDS (wide: fdsweep::kGeoLdsDoubles)

// Geometry row/layer doubles a table of this grid keeps in LDS, and what
// each build holds.
inline long long fd_geo_lds_need(int ny, int nz) {
    return fdsweep::geo_rows_doubles(ny) + fdsweep::geo_layers_doubles(nz);
}
inline long long fd_geo_lds_capacity(int variant) {
    return variant == FD_NARROW ? kNarrowGeoLds : fdsweep::kGeoLdsDoubles;
}

// Size of the device array behind fdgpu_solve_noalloc (8 MiB: one table whose
// largest face has up to about 130,000 nodes).
#define FD_NOALLOC_SCRATCH_BYTES (8u << 20)
// Device address of that array on the current device (hipGetSymbolAddress).
extern "C" hipError_t fd_noalloc_scratch(void **ptr);
// One table by one wide workgroup, descriptor and scratch layout by value.
extern "C" hipError_t fd_launch_one(const fdsweep::TableJob *J, const fdsweep::Scratch *W, hipStream_t stream);

// The persistent sweep over njobs tables on nslots workgroups; table i <
// nteam is solved by a team of d_team_first[i + 1] - d_team_first[i]
// workgroups (nteam 0: every table by one; the narrow build has no teams).
extern "C" hipError_t fd_launch_sweep(int variant, const fdsweep::TableJob *d_jobs, int njobs,
                                      const fdsweep::Scratch *d_slots, int nslots, int *d_queue,
                                      fdsweep::TeamCtl *d_teams, const int *d_team_first, int nteam,
                                      hipStream_t stream);
// Kernel address of a sweep build, for occupancy queries.
extern "C" const void *fd_sweep_kernel_ptr(int variant);
// Sets the table queue counter to 0 and clears the control words of nteam
// team blocks (a kernel, so that it can sit in a graph).
extern "C" hipError_t fd_launch_reset(int *d_queue, fdsweep::TeamCtl *d_teams, int nteam, hipStream_t stream);
// Fills the geometry table of one grid (p->sinq is a device pointer).
extern "C" hipError_t fd_launch_geo(const fdsweep::GeoParams *p, double *d_out, hipStream_t stream);
extern "C" hipError_t fd_launch_slowness(const float *d_vel, float *d_slow, long long n, hipStream_t stream);
extern "C" hipError_t fd_launch_serve(fdsweep::ServeCtl *ctl, const fdsweep::ServeJob *ring, unsigned long long *done,
                                      int cap, const fdsweep::Scratch *d_slots, int nslots, fdsweep::ServeDev *dev,
                                      long long idle_ticks, hipStream_t stream);
