// loc_kernel.hip -- gfx950 kernels of sphfdloc's grid-search location
// (src/sphfdloc.c:752-1052, find_time 1305-1345) over resident tables.
//
// One search pass of a batch of events is seven launches on one stream:
//   loc_init_kernel     per event: clear the running minimum
//   loc_nodes_kernel    nodes of layers kmin.. across the lanes, one group of
//                       the batch's events in turn over their usable picks
//                       only: a pick's table value is one coalesced global
//                       load (eight in flight per wave); the blocks of a die
//                       that run together share node tiles, so the events of
//                       all groups re-read the same lines from its L2
//   loc_commit_kernel   per event: decode the winning trial, evaluate it again
//                       (residuals and computed times of every pick, the mean
//                       residual), move the hypocentre
//   loc_refine_kernel   refinement points across the lanes, one grid row of
//                       blocks per event: trilinear times from the tables
//   (commit, refine, commit again for the second refinement)
// Every block folds its best trial into the event's key with one 64-bit
// atomic minimum; the key orders by deviation, then by scan index, so the
// result is the reference's sequential first minimum whatever the order in
// which blocks run.
//
// Compiled with -ffp-contract=off: every multiply and add rounds separately,
// as in the reference's x86-64 build.
#include <hip/hip_runtime.h>

#include "loc_core.h"

namespace {

constexpr int kThreads = 256;
constexpr int kRowsPerLaunch = 32768;   // grid rows of one launch (the y dimension stops at 65535)

// Block-wide minimum of the lanes' keys, folded into *dst.
__device__ __forceinline__ void fold_min(unsigned long long key, unsigned long long *dst) {
    __shared__ unsigned long long part[kThreads / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)key, off);
        const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(key >> 32), off);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        key = o < key ? o : key;
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; w++) key = part[w] < key ? part[w] : key;
        // the key only moves down: skip the atomic when it cannot lower it
        if (key != locdev::kNone &&
            key < __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(dst, key);
    }
    __syncthreads();
}

__global__ void loc_init_kernel(const int *__restrict__ active, int nactive, LocState *state) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nactive) return;
    LocState &s = state[active[a]];
    s.key = locdev::kNone;
    s.stdmin = __builtin_inff();   // (float)DBL_MAX
    s.avrmin = 0.0f;
    for (int q = 0; q < 3; q++) {
        s.node[q] = 0;
        s.e[q] = 0.0;
    }
}

// Sums of the usable picks [first, first + count) at node n, in pick order;
// the table loads of eight picks are issued before their arithmetic.
template <bool SCALED>
__device__ __forceinline__ void node_sums(const LocGrid &g, const LocSPick *__restrict__ sp,
                                          const unsigned char *__restrict__ sc, int first, int count, long long n,
                                          float &avres, float &rsq) {
    constexpr int U = 8;
    const int end = first + count;
    int p = first;
    for (; p + U <= end; p += U) {
        float t[U];
#pragma unroll
        for (int u = 0; u < U; u++) t[u] = locdev::gload(sp[p + u].table, n);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const float obs = sp[p + u].obs;
            const float r = SCALED && sc[p + u] ? (float)((double)obs - (double)t[u] * (double)g.vpvs)
                                                : locdev::residual_unscaled(obs, t[u]);
            locdev::add2(avres, rsq, r, sp[p + u].pwt);
        }
    }
    for (; p < end; p++) {
        const float t = locdev::gload(sp[p].table, n);
        const float obs = sp[p].obs;
        const float r = SCALED && sc[p] ? (float)((double)obs - (double)t * (double)g.vpvs)
                                        : locdev::residual_unscaled(obs, t);
        locdev::add2(avres, rsq, r, sp[p].pwt);
    }
}

// Sums at the nodes n and n + 1 (n even, both in the table): one 8-byte load
// per pick serves both; s = {avres(n), rsq(n), avres(n + 1), rsq(n + 1)}.
template <bool SCALED>
__device__ __forceinline__ void pair_sums(const LocGrid &g, const LocSPick *__restrict__ sp,
                                          const unsigned char *__restrict__ sc, int first, int count, long long n,
                                          float s[4]) {
    constexpr int U = 8;
    const int end = first + count;
    int p = first;
    for (; p + U <= end; p += U) {
        locdev::float_pair t[U];
#pragma unroll
        for (int u = 0; u < U; u++) t[u] = locdev::gload2(sp[p + u].table, n);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const float obs = sp[p + u].obs, w = sp[p + u].pwt;
            const bool scl = SCALED && sc[p + u];
            const float ra = scl ? (float)((double)obs - (double)t[u].x * (double)g.vpvs)
                                 : locdev::residual_unscaled(obs, t[u].x);
            const float rb = scl ? (float)((double)obs - (double)t[u].y * (double)g.vpvs)
                                 : locdev::residual_unscaled(obs, t[u].y);
            locdev::add2(s[0], s[1], ra, w);
            locdev::add2(s[2], s[3], rb, w);
        }
    }
    for (; p < end; p++) {
        const locdev::float_pair t = locdev::gload2(sp[p].table, n);
        const float obs = sp[p].obs, w = sp[p].pwt;
        const bool scl = SCALED && sc[p];
        const float ra = scl ? (float)((double)obs - (double)t.x * (double)g.vpvs) : locdev::residual_unscaled(obs, t.x);
        const float rb = scl ? (float)((double)obs - (double)t.y * (double)g.vpvs) : locdev::residual_unscaled(obs, t.y);
        locdev::add2(s[0], s[1], ra, w);
        locdev::add2(s[2], s[3], rb, w);
    }
}

constexpr int kPairTile = 2 * kThreads;   // nodes of a node-search block: two per lane

constexpr int kXcds = 8;   // workgroups are dealt to the accelerator dies in turn

// Block b runs on die b % kXcds.  The blocks a die holds at once should share
// node tiles, so that the tables' lines of those tiles stay in its L2 while
// every event reads them: die x takes the tiles x, x + kXcds, ..., and
// ngroups consecutive blocks of a die take the same tile, each with one
// group of the events.
__global__ __launch_bounds__(kThreads) void loc_nodes_kernel(const LocGrid g, const LocSPick *__restrict__ sp,
                                                             const LocSEvent *__restrict__ sev,
                                                             const unsigned char *__restrict__ sc,
                                                             const int *__restrict__ active, int nactive,
                                                             int ngroups, LocState *state) {
    const unsigned x = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
    const long long tile = (long long)(local / ngroups) * kXcds + x;
    const int grp = (int)(local % ngroups);
    const long long nxy = (long long)g.nx * g.ny;
    const long long n0 = nxy * g.kmin, total = nxy * g.nz;
    // node pairs from an even index: a lane's pair is one aligned 8-byte load
    const long long nbase = n0 & ~1LL;
    if (nbase + tile * kPairTile >= total) return;   // the whole block: beyond the last tile
    const long long n = nbase + tile * kPairTile + 2 * (long long)threadIdx.x;
    const bool in_a = n >= n0 && n < total, in_b = n + 1 < total;   // n + 1 >= n0 always
    const int a0 = (int)((long long)grp * nactive / ngroups), a1 = (int)((long long)(grp + 1) * nactive / ngroups);
    for (int a = a0; a < a1; a++) {
        const int ev = active[a];
        const LocSEvent E = sev[ev];
        unsigned long long key = locdev::kNone;
        if (in_b) {
            float sums[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (E.scaled)
                pair_sums<true>(g, sp, sc, E.first, E.count, n, sums);
            else
                pair_sums<false>(g, sp, sc, E.first, E.count, n, sums);
            const unsigned long long kb = locdev::key_of(locdev::deviation(sums[2], sums[3], E.facs, E.den),
                                                         __builtin_inff(), (unsigned)(n + 1 - n0));
            key = kb;
            if (in_a) {
                const unsigned long long ka = locdev::key_of(locdev::deviation(sums[0], sums[1], E.facs, E.den),
                                                             __builtin_inff(), (unsigned)(n - n0));
                key = ka < kb ? ka : kb;
            }
        } else if (in_a) {   // the grid's last node, without a partner
            float avres = 0.0f, rsq = 0.0f;
            if (E.scaled)
                node_sums<true>(g, sp, sc, E.first, E.count, n, avres, rsq);
            else
                node_sums<false>(g, sp, sc, E.first, E.count, n, avres, rsq);
            key = locdev::key_of(locdev::deviation(avres, rsq, E.facs, E.den), __builtin_inff(), (unsigned)(n - n0));
        }
        fold_min(key, &state[ev].key);
    }
}

__global__ __launch_bounds__(kThreads) void loc_refine_kernel(const LocGrid g, int second,
                                                              const LocSPick *__restrict__ sp,
                                                              const LocSEvent *__restrict__ sev,
                                                              const unsigned char *__restrict__ sc,
                                                              const int *__restrict__ active, LocState *state) {
    const int ev = active[blockIdx.y];
    const LocSEvent E = sev[ev];
    const LocState &st = state[ev];
    const locdev::Box b = second ? locdev::box_second(g, st.e) : locdev::box_first(g, st.node);
    const unsigned npts = locdev::box_points(b);
    if ((unsigned)blockIdx.x * kThreads >= npts) return;   // the whole block: beyond this event's box
    const unsigned q = blockIdx.x * kThreads + threadIdx.x;
    unsigned long long key = locdev::kNone;
    if (q < npts) {
        double pt[3];
        locdev::box_point(b, q, pt);
        const locdev::Cell c = locdev::cell_of(g, pt[0], pt[1], pt[2]);
        unsigned off[8];
        locdev::corner_offsets(c, g.nx, (long long)g.nx * g.ny, off);
        float avres = 0.0f, rsq = 0.0f;
#pragma unroll 2
        for (int p = E.first; p < E.first + E.count; p++) {
            double t = locdev::interp_b(locdev::table_rsrc(sp[p].table, g.tbytes), c, off);
            if (E.scaled && sc[p]) t *= (double)g.vpvs;
            locdev::add2(avres, rsq, (float)((double)sp[p].obs - t), sp[p].pwt);
        }
        key = locdev::key_of(locdev::deviation(avres, rsq, E.facs, E.den), st.stdmin, q);
    }
    fold_min(key, &state[ev].key);
}

// Evaluates the winning trial again for every pick (sphfdloc.c:793-803):
// the same functions in the same order give the same bits as the search.
__global__ void loc_commit_kernel(const LocGrid g, int stage, const LocPick *__restrict__ picks,
                                  const LocEvent *__restrict__ events, const int *__restrict__ active, int nactive,
                                  LocState *state, float *__restrict__ res, float *__restrict__ tps) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nactive) return;
    const int ev = active[a];
    const LocEvent E = events[ev];
    LocState st = state[ev];
    const unsigned long long key = st.key;
    const long long sy = g.nx, sz = (long long)g.nx * g.ny;
    long long node_n = 0;
    locdev::Cell c;
    if (stage == 0) {
        int nd[3] = {0, 0, 0};
        if (key != locdev::kNone) {
            node_n = sz * g.kmin + (long long)(unsigned)key;
            nd[0] = (int)(node_n % g.nx);
            nd[1] = (int)((node_n / g.nx) % g.ny);
            nd[2] = (int)(node_n / sz);
        }
        const double org[3] = {g.x0, g.y0, g.z0}, sp[3] = {g.df, g.dq, g.h};
        for (int q = 0; q < 3; q++) {
            st.node[q] = nd[q];
            st.e[q] = org[q] + nd[q] * sp[q];
        }
    } else if (key != locdev::kNone) {
        const locdev::Box b = stage == 2 ? locdev::box_second(g, st.e) : locdev::box_first(g, st.node);
        double pt[3];
        locdev::box_point(b, (unsigned)key, pt);
        for (int q = 0; q < 3; q++) st.e[q] = pt[q];
        c = locdev::cell_of(g, pt[0], pt[1], pt[2]);
    }
    if (key != locdev::kNone) {
        locdev::Sums s = {0.0f, 0.0f, 0.0f};
        for (int p = E.first; p < E.first + E.npick; p++) {
            const LocPick P = picks[p];
            const double t = stage == 0 ? (double)P.table[node_n] : locdev::interp(P.table, c, sy, sz);
            const float r = locdev::residual(g, P, t);
            if (P.good) locdev::add(s, r, P.pwt);
            res[p] = r;
            tps[p] = P.obs - r;
        }
        st.stdmin = locdev::deviation(s, E.ngood);
        st.avrmin = s.avres / s.facs;
    }
    st.key = locdev::kNone;
    state[ev] = st;
}

}  // namespace

extern "C" hipError_t fd_launch_loc_pass(const LocGrid *grid, const LocPick *d_picks, const LocEvent *d_events,
                                         const LocSPick *d_spicks, const LocSEvent *d_sevents,
                                         const unsigned char *d_sscale, const int *d_active, int nactive,
                                         LocState *d_state, float *d_res, float *d_tps, hipEvent_t node0,
                                         hipEvent_t node1, hipStream_t stream) {
    if (nactive <= 0) return hipSuccess;
    const LocGrid g = *grid;
    const long long nodes = (long long)g.nx * g.ny * (g.nz - g.kmin);
    const int small = (nactive + 63) / 64;
    hipLaunchKernelGGL(loc_init_kernel, dim3(small), dim3(64), 0, stream, d_active, nactive, d_state);
    if (node0) (void)hipEventRecord(node0, stream);
    if (nodes > 0) {
        // at most 16 groups of at least 8 events: a die then holds about
        // 8 x 32 CUs / ngroups tiles at once, their table lines within its L2
        const int ngroups = nactive >= 128 ? 16 : nactive >= 8 ? nactive / 8 : 1;
        const long long nbase = (long long)g.nx * g.ny * g.kmin & ~1LL;
        const long long tiles = ((long long)g.nx * g.ny * g.nz - nbase + kPairTile - 1) / kPairTile;
        const long long per_die = (tiles + kXcds - 1) / kXcds;
        hipLaunchKernelGGL(loc_nodes_kernel, dim3((unsigned)(per_die * ngroups * kXcds)), dim3(kThreads), 0, stream, g,
                           d_spicks, d_sevents, d_sscale, d_active, nactive, ngroups, d_state);
    }
    if (node1) (void)hipEventRecord(node1, stream);
    hipLaunchKernelGGL(loc_commit_kernel, dim3(small), dim3(64), 0, stream, g, 0, d_picks, d_events, d_active,
                       nactive, d_state, d_res, d_tps);
    for (int stage = 1; stage <= 2; stage++) {
        const long long side = 2LL * (stage == 1 ? g.ndiv : g.ndiv2) + 1;
        const long long pts = side * side * side;
        // one grid row per event, at most kRowsPerLaunch rows per launch
        for (int first = 0; first < nactive; first += kRowsPerLaunch) {
            const int rows = nactive - first < kRowsPerLaunch ? nactive - first : kRowsPerLaunch;
            hipLaunchKernelGGL(loc_refine_kernel, dim3((unsigned)((pts + kThreads - 1) / kThreads), (unsigned)rows),
                               dim3(kThreads), 0, stream, g, stage == 2 ? 1 : 0, d_spicks, d_sevents, d_sscale,
                               d_active + first, d_state);
        }
        hipLaunchKernelGGL(loc_commit_kernel, dim3(small), dim3(64), 0, stream, g, stage, d_picks, d_events,
                           d_active, nactive, d_state, d_res, d_tps);
    }
    return hipGetLastError();
}
