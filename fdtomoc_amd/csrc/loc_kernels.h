// loc_kernels.h -- device-side records and launch wrapper of the grid-search
// location of FDtomoC's sphfdloc (src/sphfdloc.c:752-1052, find_time
// 1305-1345) over resident travel-time tables, shared by the kernels
// (loc_kernel.hip) and the host side (loc.cpp).
#pragma once

#include <hip/hip_runtime.h>

// The search grid and parameters in sphfdloc's units after its set-up
// (sphfdloc.c:382-393): longitude and geocentric colatitude in radians,
// depth in km.
struct LocGrid {
    int nx, ny, nz;
    int kmin, ndiv, ndiv2;
    float vpvs;
    unsigned tbytes;   // bytes of one table (nx*ny*nz < 2^30)
    double x0, y0, z0, df, dq, h;
};

// One pick: its table (nz*ny*nx floats, x fastest), its time after the
// event's reference time, its weight, whether it counts in the sums, and
// whether its time is vpvs times its table (ivs == 0 and an S pick).
struct LocPick {
    const float *table;
    float obs, pwt;
    int good, scale;
};

// An event's picks: [first, first + npick) of the pick array; ngood is the
// count the variance divides by in this pass.
struct LocEvent {
    int first, npick, ngood, pad;
};

// The search kernels' view of one pass: each event's usable picks only,
// contiguous, and the terms of its variance that no trial changes.
struct LocSPick {
    const float *table;
    float obs, pwt;
};
struct LocSEvent {
    int first, count;   // [first, first + count) of the usable-pick array
    int scaled;         // some usable pick's time is vpvs times its table
    float facs;         // sum of the usable picks' weights in pick order (sphfdloc.c:781)
    float den;          // (ngood - 4) * (facs / ngood) (sphfdloc.c:786-788)
    int pad;
};

// Running state of one event's search pass: the best trial of the running
// stage as (deviation bits << 32 | scan index), the minimum so far, the
// hypocentre and the node of the node search.
struct LocState {
    unsigned long long key;
    float stdmin, avrmin;
    double e[3];
    int node[3];
    int pad;
};

// One search pass for the events listed in d_active: the node search from
// layer kmin, then the two refinements; d_state, d_res and d_tps receive
// each event's minimum, hypocentre, and residuals and computed times at it.
// node0/node1 (or null) are recorded around the node search.
// d_spicks/d_sevents (indexed by event like d_events) and d_sscale (one flag
// per usable pick: time = vpvs * table) are the pass's usable picks.
extern "C" hipError_t fd_launch_loc_pass(const LocGrid *grid, const LocPick *d_picks, const LocEvent *d_events,
                                         const LocSPick *d_spicks, const LocSEvent *d_sevents,
                                         const unsigned char *d_sscale, const int *d_active, int nactive,
                                         LocState *d_state, float *d_res, float *d_tps, hipEvent_t node0,
                                         hipEvent_t node1, hipStream_t stream);
