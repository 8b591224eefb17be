// sweep_solver.h -- one workgroup solves one travel-time table: the whole
// expanding-box sweep of FDtomoC src/sphfd.c (first pass plus head-wave reverse
// passes, sphfd.c:1008-5678), written bulk-synchronously against a small
// executor interface X:
//
//   int  X::tid() const, X::nthr() const   thread index / count in the group
//   void X::sync()                         group barrier with memory ordering
//   int  X::atomic_add(int *p, int v)      returns the old value
//   int  X::atomic_exch(int *p, int v)     returns the old value
//   unsigned X::atomic_sub_u(unsigned *p, unsigned v)  returns the old value
//   unsigned X::atomic_or(unsigned*,unsigned), X::atomic_and(...)
//   long long X::clock() const             a monotonic tick counter (timing only)
//   unsigned long long X::atomic_max_u64(...), unsigned X::atomic_cas(...)
//   void X::fence_release(), X::fence_acquire()   agent-scope fences (teams)
//   void X::wait_stores()                  this thread's stores have completed
//   int  X::load_relaxed(const int*), void X::store_relaxed(int*, int),
//   void X::store_release(int*, int)
//   void X::pause()                        back off inside a wait loop
//
// On the GPU X wraps threadIdx/__syncthreads/atomics (sweep_kernel.hip).
//
// Exactness.  Inside a face the reference updates nodes one by one in
// ascending (inner-layer time, fill index) order; a node reads the new values
// of in-face neighbours that precede it and the pre-face values of those that
// follow it.  Two schedules reproduce that bit for bit:
//
//   LEVEL  (faces holding uncomputed nodes, i.e. the expanding first pass)
//     A stable radix sort gives every node its rank.  A node becomes ready when
//     all its preceding in-face neighbours are final (pending counters).  Ready
//     nodes never neighbour each other, so a round updates them in parallel.
//     fdsph2d also reads t[0]/t[4] left by whichever node ran before it in the
//     global order (sphfd.c:6116-6121); a node that runs a 2D trial without
//     having set them itself walks back through the rank order to the last node
//     that did, and waits a round if that node is not final yet.
//
//   CHANGE (faces whose nodes are all computed, i.e. reverse passes)
//     Every node is evaluated once from a snapshot of the pre-face values;
//     afterwards a node is re-evaluated only when a preceding neighbour changed.
//     This converges to the sequential result (it is the same DAG evaluated in
//     topological order) and needs far fewer serial rounds, because reverse
//     passes change few nodes.  If any node would need the carried t[0]/t[4],
//     the face is restored from the snapshot and redone with LEVEL.
#pragma once

#include "sweep_core.h"

namespace fdsweep {

// --------------------------------------------------------------- job layout

// One table, prepared by the host (geometry and source initialisation follow
// sphfd.c:415-509 and 696-975, evaluated with host libm).
struct TableJob {
    int nx, ny, nz;
    double h, dq, df, y0, z0r;
    const double *sinq;        // ny + 2 values, sin(qc(j)) at [j + 1]
    const float *slow;         // slowness at node (0,0,0) of this table
    long long s_sy, s_sz;      // slowness strides (x stride is 1)
    float *time;               // nx*ny*nz output grid, reference layout
    int init_mode;             // 0: fill 1e10, cube times s000*dist; 1: grid preset; 2: fill, explicit values
    long long src_sidx;        // slowness index of the source node (s000, sphfd.c:730)
    int ninit;
    const int *init_idx;
    const double *init_dist;   // mode 0: straight-ray distances (sphfd.c:727-749)
    const float *init_val;     // mode 2: box-file values (sphfd.c:763)
    int box[3][2];             // x1 x2 / y1 y2 / z1 z2 after source setup
    int dflag[6];              // dx1 dx2 dy1 dy2 dz1 dz2
    int grow[6];               // kminus kplus iplus iminus jminus jplus
    int reverse, headpref, maxrad, radius0, srcwall;
    int *stats;                // kStat* counters (device)
    int *mod_stamp;            // nx*ny*nz: face counter of each node's last change (NULL:
                               // the solving slot's Scratch::mstamp, or not kept when that
                               // is NULL too, and then skip_ok must be 0)
    int *last_full;            // 6 * maxdim: face counter at the start of the last
                               // full-plane face per (face type, plane) (NULL: the
                               // slot's Scratch::lastf)
    int maxdim;                // max(nx, ny, nz)
    int skip_ok;               // skipping unchanged reverse-pass nodes is exact here
    int change_budget;         // CHANGE evaluations per face node before a face goes to
                               // LEVEL (0: never); a schedule choice, results are identical
    const double *geo;         // geometry tables of this grid (geo_doubles(ny, nz)), or
    double *geo_fill;          // if set, a buffer the workgroup fills first and then uses
    int wave;                  // passes run as one dataflow wavefront (sweep_wave.h) instead of
                               // face by face: bit0 the first pass, bit1 reverse passes (the
                               // host sets it only where the slot holds the wavefront state)
};

enum {
    kStatPasses = 0, kStatRestarts, kStatFacesLevel, kStatFacesChange, kStatRoundsLevel,
    kStatRoundsChange, kStatFallbacks, kStatEvals, kStatDefers, kStatError, kStatLastWall,
    kStatHeadw0,  // 7 entries (slots 0..6) of the first pass
    kStatSkips = kStatHeadw0 + 7,  // CHANGE first-round evaluations skipped (inputs unchanged)
    // time spent by the table, in X::clock() ticks (10 ns on the GPU)
    kStatTimeGather, kStatTimeSort, kStatTimeLevel, kStatTimeChange, kStatTimeInit,
    // LEVEL rounds, as seen by lane 0 (which always holds the first ready node):
    // ticks from the round's start to its evaluation, ticks of the evaluation
    kStatTimeLevelPre, kStatTimeLevelEval,
    kStatWalks,         // LEVEL: rank-order steps walked back to find carried t[0]/t[4]
    kStatTimeChange1,   // CHANGE: ticks of the first round (every node of the face)
    // wavefront passes (sweep_wave.h)
    kStatWavePasses,    // passes run as a wavefront
    kStatRoundsWave,    // their rounds
    kStatWaveSeeds,     // nodes ready at the start of a pass (no predecessors)
    kStatWaveCarry,     // nodes that waited for the carried t[0]/t[4] of their predecessor
    kStatWaveScans,     // face scans for such predecessors
    kStatTimeWave,      // ticks spent in wavefront passes
    kStatTimeWaveGroup, // of which grouping each round's nodes by face type
    kStatTimeWaveEval,  // evaluations and queueing (phase 3)
    kStatTimeWaveCarry, // carried-register requests (phase 4)
    kStatTeam,          // workgroups that solved the table (a team's size, else 1)
    kStatWaveFail,      // why a wavefront pass failed: 1 list capacity, 2 a key not final at a
                        // request's face scan, 3 requests that never resolved, 4 round limit,
                        // 5 closing list capacity, 6 evaluation count, 7 a member waited too long
    kStatCount
};

// Per-slot scratch, sized for the largest face of any table in the launch.
struct Scratch {
    float *O;            // pre-face values of the face nodes (fill order)
    unsigned *key;       // ordered key of the inner-layer time (fill order)
    unsigned *key2;      // sort ping-pong
    int *val, *val2;     // sort ping-pong payload (fill index)
    int *rank;           // fill index -> rank
    int *ord;            // rank -> fill index
    int *pend;           // LEVEL, faces too large for LDS: byte-packed pending counters
    int *stat;           // per rank: 0 not final, else (round+1)<<2 | 1 set t0/t4, 2 passed them on
    float *tout0, *tout4;  // per rank (LEVEL) or fill index (CHANGE)
    int *list0, *list1, *list2;  // work lists
    int *stamp;          // CHANGE: last round a node was queued
    unsigned char *flag; // CHANGE: bit0 head wave, bit1 used carried t0/t4, bit2 sets t0/t4, bit3 evaluated;
                         // LEVEL, faces too large for LDS: successor masks
    int *mstamp;         // per-node change stamps of the table being solved (sized for the
                         // largest table of the launch), or NULL
    int *lastf;          // its last full-plane faces (6 * largest dimension), or NULL
    // wavefront passes (sweep_wave.h), sized for the largest table of the launch
    unsigned *wst;       // per node: bit0 final, bit1 set t[0]/t[4], bits 2.. round (final) or
                         // last round queued (not final); rounds never repeat on a slot
    float *wt0, *wt4;    // per node: t[0]/t[4] as the node left them (or its resolved carry-in)
    int *wcnt;           // per node: earlier-face predecessors still to finalise (bits 0-7) and
                         // in-face predecessors finalised (bits 8-15)
    int *wtmp, *wready;  // the lists of the rounds: one evaluated, one filled for the next round
    int *wclose;         // teams: per round, nodes whose earlier-face predecessors all finished
    long long wcap;      // capacity of each of the three lists
    long long wn;        // nodes wst, wt0 and wt4 hold (0: no wavefront state)
    unsigned *wround;    // the slot's next round number (one word; starts at 1 over zeroed
                         // status words)
};

// The stamps and last full-plane faces a table's solve uses: its own, else
// the solving slot's (they live only for the duration of one solve).
FD_INLINE int *stamps_of(const TableJob &J, const Scratch &W) { return J.mod_stamp ? J.mod_stamp : W.mstamp; }
FD_INLINE int *lastfull_of(const TableJob &J, const Scratch &W) { return J.last_full ? J.last_full : W.lastf; }

// LDS scratch of a workgroup: radix histograms during a sort, then the LEVEL
// schedule's byte-packed pending counters and successor masks (2 bytes per
// face node: faces up to 52,000 nodes with the wide build's 26,000 words,
// 26,800 with the narrow build's 13,400; larger faces keep them in global
// scratch), or the CHANGE schedule's
// queued-follower bits (one bit per node).  LW bounds the words the schedules
// may use for per-node state (LW <= AW; host test builds set it to 0 to run
// every face through the global-scratch paths).  Beside it, the geometry row
// and layer tables of the table being solved: geo_rows_doubles(ny) +
// geo_layers_doubles(nz) doubles, at most kGeoLds (the host sends larger
// grids to a build that holds them, or refuses them).
constexpr int kArenaWords = 26000;
constexpr int kGeoLdsDoubles = 6800;

template <int AW, int GL = kGeoLdsDoubles, int LW = AW>
struct SharedT {
    static constexpr int kArena = AW;
    static constexpr int kNodeWords = LW;
    static_assert(LW <= AW, "per-node state must fit the arena");
    static constexpr int kGeoLds = GL;
    double geol[GL];     // geometry row and layer tables
    int n[3];            // list sizes
    int processed;
    int headw[7];
    unsigned kor, kand;
    int flag;            // face flags: bit0 uncomputed node present, bit1 NaN key, bit2 fall back
    int best;            // CHANGE: fill index of the last node in face order
    float cin0, cin4;    // carried t[0]/t[4] entering the face
    int job;
    int face_ctr;        // faces swept so far in this table (stamps start at 1)
    unsigned long long wbest;  // wavefront: result of a face scan
    int wc[32];          // wavefront: counters and flags
    int st[kStatCount];  // counters of the current table, flushed to TableJob::stats at the end
    int arena[AW];
};

// Workgroup state of the one-workgroup-per-CU kernel (and of host builds).
using Shared = SharedT<kArenaWords>;

// ---------------------------------------------------------------- helpers

FD_INLINE unsigned order_key(float v) {
    unsigned u;
    __builtin_memcpy(&u, &v, 4);
    if (u == 0x80000000u) u = 0;   // -0 sorts with +0 (the reference compares floats)
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

FD_INLINE bool key_less(unsigned ka, int a, unsigned kb, int b) { return ka < kb || (ka == kb && a < b); }

struct FaceGeom {
    int face, F, wa, wb, a0, b0, fixed;
};

template <int F>
FD_INLINE void face_node(const FaceGeom &fg, int p, int &i, int &j, int &k, int &a, int &b) {
    a = p % fg.wa;
    b = p / fg.wa;
    const int A = fg.a0 + a, B = fg.b0 + b;
    if (F == TOP || F == BOTTOM) { i = A; j = B; k = fg.fixed; }
    else if (F == EAST || F == WEST) { i = fg.fixed; j = A; k = B; }
    else { i = A; j = fg.fixed; k = B; }
}

// Offsets of an in-face neighbour along the face's (a, b) axes.
template <int F>
FD_INLINE void inface(int dx, int dy, int dz, int &dn, int &da, int &db) {
    if (F == TOP || F == BOTTOM) { dn = dz; da = dx; db = dy; }
    else if (F == EAST || F == WEST) { dn = dx; da = dy; db = dz; }
    else { dn = dy; da = dx; db = dz; }
}

// Loads a node's 3x3x2 neighbourhood.  CHG: in-face neighbours that follow the
// node in face order contribute their pre-face value from the snapshot O, the
// others the grid (tentative or final).  LEVEL reads the grid only: there the
// followers have not been touched yet.  Coordinates are clamped into the grid;
// trials never use a clamped entry because their bounds checks fail first.
template <int F, bool CHG>
FD_INLINE void load_nbhd(const TableJob &J, const Geo &G, const FaceGeom &fg, const float *O, const unsigned *key,
                         int p, int a, int b, int i, int j, int k, Nbhd &N) {
    using S = FaceSpec<F>;
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    // every load below is unconditional and independent, so all of them are in
    // flight together; the CHANGE selection happens afterwards on registers
    unsigned kq[9];
    float oq[9];
#pragma unroll
    for (int pl = 0; pl < 2; pl++)
#pragma unroll
        for (int db = -1; db <= 1; db++)
#pragma unroll
            for (int da = -1; da <= 1; da++) {
                int o[3] = {0, 0, 0};
                o[S::nax] = pl ? S::ndir : 0;
                o[S::a_ax] = da;
                o[S::b_ax] = db;
                int ci = i + o[0], cj = j + o[1], ck = k + o[2];
                ci = ci < 0 ? 0 : (ci >= J.nx ? J.nx - 1 : ci);
                cj = cj < 0 ? 0 : (cj >= J.ny ? J.ny - 1 : cj);
                ck = ck < 0 ? 0 : (ck >= J.nz ? J.nz - 1 : ck);
                const int q = pl * 9 + (db + 1) * 3 + (da + 1);
                N.t[q] = J.time[(long long)ck * nxy + (long long)cj * nx + ci];
                N.s[q] = J.slow[(long long)ck * J.s_sz + (long long)cj * J.s_sy + ci];
                if (CHG && pl == 0) {
                    int aa = a + da, bb = b + db;
                    aa = aa < 0 ? 0 : (aa >= fg.wa ? fg.wa - 1 : aa);
                    bb = bb < 0 ? 0 : (bb >= fg.wb ? fg.wb - 1 : bb);
                    kq[q] = key[bb * fg.wa + aa];
                    oq[q] = O[bb * fg.wa + aa];
                }
            }
    N.sq[0] = G.sinq[j];
    N.sq[1] = G.sinq[j + 1];
    N.sq[2] = G.sinq[j + 2];
    if (CHG) {
        const unsigned kp = kq[4];
#pragma unroll
        for (int db = -1; db <= 1; db++)
#pragma unroll
            for (int da = -1; da <= 1; da++) {
                if (!(da | db)) continue;
                const int q = (db + 1) * 3 + (da + 1);
                const int aa = a + da, bb = b + db;
                const bool inside = aa >= 0 && bb >= 0 && aa < fg.wa && bb < fg.wb;
                const int fq = bb * fg.wa + aa;
                // a follower (later in face order) contributes its pre-face value
                if (inside && !key_less(kq[q], fq, kp, p)) N.t[q] = oq[q];
            }
    }
}

// Latest change stamp over a node's 3x3x2 neighbourhood (its every input).
template <int F>
FD_INLINE int max_stamp(const TableJob &J, const int *ms, int i, int j, int k) {
    using S = FaceSpec<F>;
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    int m = 0;
#pragma unroll
    for (int pl = 0; pl < 2; pl++)
#pragma unroll
        for (int db = -1; db <= 1; db++)
#pragma unroll
            for (int da = -1; da <= 1; da++) {
                int o[3] = {0, 0, 0};
                o[S::nax] = pl ? S::ndir : 0;
                o[S::a_ax] = da;
                o[S::b_ax] = db;
                int ci = i + o[0], cj = j + o[1], ck = k + o[2];
                ci = ci < 0 ? 0 : (ci >= J.nx ? J.nx - 1 : ci);
                cj = cj < 0 ? 0 : (cj >= J.ny ? J.ny - 1 : cj);
                ck = ck < 0 ? 0 : (ck >= J.nz ? J.nz - 1 : ck);
                const int v = ms[(long long)ck * nxy + (long long)cj * nx + ci];
                m = v > m ? v : m;
            }
    return m;
}

// ------------------------------------------------------------ radix sort

// Stable LSD sort of F (key, val) pairs by key, 4-bit digits, skipping digits
// on which all keys agree.  Returns true if the result is in (key2, val2).
template <class X, class SH>
FD_INLINE bool radix_sort(X &x, SH &sh, unsigned *k0, int *v0, unsigned *k1, int *v1, int F) {
    const int T = x.nthr(), tid = x.tid();
    const int chunk = (F + T - 1) / T;
    const int lo = tid * chunk < F ? tid * chunk : F;
    const int hi = lo + chunk < F ? lo + chunk : F;
    int *hist = sh.arena;              // 16 * T digit-major counters
    int *part = sh.arena + 16 * T;     // T partial sums
    int *csum = part + T;              // their sums in chunks of 16
    bool swapped = false;
    for (int d = 0; d < 8; d++) {
        const unsigned vary = (sh.kor ^ sh.kand) >> (4 * d);
        if ((vary & 15u) == 0) continue;
        // per-thread digit counts live in this thread's column of hist
        for (int q = 0; q < 16; q++) hist[q * T + tid] = 0;
        for (int e = lo; e < hi; e++) hist[((k0[e] >> (4 * d)) & 15u) * T + tid]++;
        x.sync();
        // exclusive scan over 16*T counters in digit-major order: each thread
        // owns 16 consecutive counters
        int s = 0;
        for (int q = 0; q < 16; q++) s += hist[tid * 16 + q];
        part[tid] = s;
        x.sync();
        // exclusive scan of the T partial sums in chunks of 16: chunk sums,
        // a scan of the chunk sums, then each chunk's prefixes (a serial depth
        // of 16 + T/16 + 16 adds instead of T)
        const int nch = (T + 15) / 16;
        if (tid < nch) {
            int a = 0;
            for (int q = 16 * tid; q < 16 * tid + 16 && q < T; q++) a += part[q];
            csum[tid] = a;
        }
        x.sync();
        if (tid == 0) {
            int run = 0;
            for (int c = 0; c < nch; c++) {
                const int v = csum[c];
                csum[c] = run;
                run += v;
            }
        }
        x.sync();
        if (tid < nch) {
            int run = csum[tid];
            for (int q = 16 * tid; q < 16 * tid + 16 && q < T; q++) {
                const int v = part[q];
                part[q] = run;
                run += v;
            }
        }
        x.sync();
        int run = part[tid];
        for (int q = 0; q < 16; q++) {
            const int v = hist[tid * 16 + q];
            hist[tid * 16 + q] = run;
            run += v;
        }
        x.sync();
        for (int e = lo; e < hi; e++) {
            const unsigned kk = k0[e];
            const int dg = (kk >> (4 * d)) & 15u;
            const int pos = hist[dg * T + tid]++;
            k1[pos] = kk;
            v1[pos] = v0[e];
        }
        x.sync();
        unsigned *tk = k0; k0 = k1; k1 = tk;
        int *tv = v0; v0 = v1; v1 = tv;
        swapped = !swapped;
    }
    return swapped;
}

// ------------------------------------------------------------- one face

// In-face neighbour directions of the successor masks.
FD_INLINE int dir_da(int d) { return d == 0 || d == 3 || d == 5 ? -1 : (d == 1 || d == 6 ? 0 : 1); }
FD_INLINE int dir_db(int d) { return d < 3 ? -1 : (d < 5 ? 0 : 1); }

template <int F, class X, class SH>
FD_INLINE void face_level(X &x, SH &sh, const TableJob &J, const Scratch &W, Geo &G, const FaceGeom &fg,
                          const int *ord, int fcur) {
    const int T = x.nthr(), tid = x.tid();
    const int Fn = fg.F, wa = fg.wa, wb = fg.wb;
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    for (int r = tid; r < Fn; r += T) W.rank[ord[r]] = r;
    // pending counters, one byte per node packed four to a word, and successor
    // masks (bit d: the neighbour in direction d follows the node)
    const int nwords = (Fn + 3) / 4;
    const bool in_lds = (long long)nwords * 4 + Fn <= (long long)SH::kNodeWords * 4;
    unsigned *pend = in_lds ? (unsigned *)sh.arena : (unsigned *)W.pend;
    unsigned char *succ = in_lds ? (unsigned char *)(sh.arena + nwords) : W.flag;
    int *lists[3] = {W.list0, W.list1, W.list2};
    if (tid == 0) { sh.n[0] = 0; sh.n[1] = 0; sh.n[2] = 0; sh.processed = 0; }
    x.sync();
    for (int w = tid; w < nwords; w += T) {
        unsigned word = 0;
        for (int e = 0; e < 4; e++) {
            const int p = 4 * w + e;
            if (p >= Fn) break;
            const int a = p % wa, b = p / wa, rp = W.rank[p];
            int c = 0;
            unsigned m = 0;
            for (int d = 0; d < 8; d++) {
                const int aa = a + dir_da(d), bb = b + dir_db(d);
                if (aa < 0 || bb < 0 || aa >= wa || bb >= wb) continue;
                if (W.rank[bb * wa + aa] < rp) c++;
                else m |= 1u << d;
            }
            word |= (unsigned)c << (8 * e);
            succ[p] = (unsigned char)m;
            W.stat[rp] = 0;
            if (c == 0) lists[0][x.atomic_add(&sh.n[0], 1)] = p;
        }
        pend[w] = word;
    }
    x.sync();
    // One barrier per round: round R evaluates list R%3, fills list (R+1)%3 and
    // clears the counter of list (R+2)%3, which no one touches during round R.
    int R = 0;
    for (;; R++) {
        const int lc = R % 3, ln = (R + 1) % 3, lz = (R + 2) % 3;
        const int ncur = sh.n[lc];
        if (ncur == 0) break;
        if (R > 2 * Fn + 8) {   // every round finalises the lowest open rank
            if (tid == 0) sh.st[kStatError] |= 8;
            break;
        }
        if (tid == 0) sh.n[lz] = 0;
        const int *cur = lists[lc];
        int *nxt = lists[ln];
        int ndone = 0, ndefer = 0, nwalk = 0;
        const long long tr0 = tid == 0 ? x.clock() : 0;
        for (int it = tid; it < ncur; it += T) {
            const int p = cur[it];
            int i, j, k, a, b;
            face_node<F>(fg, p, i, j, k, a, b);
            Nbhd N;
            load_nbhd<F, false>(J, G, fg, W.O, W.key, p, a, b, i, j, k, N);
            const long long node = (long long)k * nxy + (long long)j * nx + i;
            const float told = W.O[p];
            const int rp = W.rank[p];
            float tin0 = sh.cin0, tin4 = sh.cin4;
            if (needs_carry<F>(N, G, i, j, k, told)) {
                // the last earlier node that set t[0]/t[4]; statuses written in
                // this round are not visible yet and count as open
                int r = rp - 1;
                bool open = false;
                while (r >= 0) {
                    const int v = W.stat[r];
                    nwalk++;
                    if (v == 0 || (v >> 2) > R) { open = true; break; }
                    if ((v & 3) == 1) break;
                    r--;
                }
                if (open) {
                    nxt[x.atomic_add(&sh.n[ln], 1)] = p;
                    ndefer++;
                    continue;
                }
                if (r >= 0) {
                    tin0 = W.tout0[r];
                    tin4 = W.tout4[r];
                }
            }
            NodeOut o;
            const long long te0 = it == 0 ? x.clock() : 0;
            eval_node<F>(N, G, i, j, k, told, tin0, tin4, o);
            if (it == 0) {
                sh.st[kStatTimeLevelPre] += (int)(te0 - tr0);
                sh.st[kStatTimeLevelEval] += (int)(x.clock() - te0);
            }
            if (o.stored) {
                J.time[node] = o.tnew;
                if (int *ms = stamps_of(J, W)) ms[node] = fcur;
                if (o.head) x.atomic_add(&sh.headw[face_slot(F)], 1);
            }
            W.tout0[rp] = o.tout0;
            W.tout4[rp] = o.tout4;
            W.stat[rp] = ((R + 1) << 2) | (o.sets ? 1 : 2);
            // release the successors; the byte cannot borrow, it is >= 1
            const unsigned m = succ[p];
            for (int d = 0; d < 8; d++) {
                if (!(m & (1u << d))) continue;
                const int q = (b + dir_db(d)) * wa + a + dir_da(d);
                const int shf = 8 * (q & 3);
                const unsigned old = x.atomic_sub_u(&pend[q >> 2], 1u << shf);
                if (((old >> shf) & 255u) == 1u) nxt[x.atomic_add(&sh.n[ln], 1)] = q;
            }
            ndone++;
        }
        if (ndone) {
            x.atomic_add(&sh.processed, ndone);
            x.atomic_add(&sh.st[kStatEvals], ndone);
        }
        if (ndefer) x.atomic_add(&sh.st[kStatDefers], ndefer);
        if (nwalk) x.atomic_add(&sh.st[kStatWalks], nwalk);
        if (tid == 0) sh.st[kStatRoundsLevel] += 1;
        x.sync();
    }
    // t[0]/t[4] leaving the face: those of the last node in face order that set them
    if (tid == 0) {
        int r = Fn - 1;
        while (r >= 0 && (W.stat[r] & 3) == 2) r--;
        if (r >= 0) { sh.cin0 = W.tout0[r]; sh.cin4 = W.tout4[r]; }
        if (sh.processed != Fn) sh.st[kStatError] |= 4;
    }
    x.sync();
}

// CHANGE schedule.  Returns false (face restored) if LEVEL must redo the face.
// Work list: W.list0 (nodes to evaluate), W.list2 (nodes that changed).
template <int F, class X, class SH>
FD_INLINE bool face_change(X &x, SH &sh, const TableJob &J, const Scratch &W, Geo &G, const FaceGeom &fg,
                       int fcur, int fprev) {
    const int T = x.nthr(), tid = x.tid();
    const int Fn = fg.F;
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    if (tid == 0) { sh.n[0] = 0; sh.n[2] = 0; sh.flag &= ~4; sh.best = 0; }
    // followers queued in the current round: a bit per node in LDS when the
    // face fits, else a round stamp per node in scratch
    const int nbits = (Fn + 31) / 32;
    const bool qlds = nbits <= SH::kNodeWords;
    unsigned *qbits = (unsigned *)sh.arena;
    if (!qlds)
        for (int p = tid; p < Fn; p += T) W.stamp[p] = 0;
    x.sync();
    long long spent = 0;
    const long long tc0 = x.clock();
    for (int round = 1;; round++) {
        const int nwork = (round == 1) ? Fn : sh.n[0];
        if (nwork == 0) break;
        // When changes keep cascading through the face (typically the first
        // reverse pass), every round re-evaluates most of it; LEVEL evaluates
        // each node once, so past change_budget evaluations per node the face
        // is restored and handed to LEVEL.
        if (round > 1 && J.change_budget > 0 && spent + nwork > (long long)J.change_budget * Fn) {
            if (tid == 0) sh.flag |= 4;
            x.sync();
            break;
        }
        spent += nwork;
        // the evaluation DAG has depth <= F, so F + 1 rounds always suffice
        if (round > 2 * Fn + 8) {
            if (tid == 0) { sh.st[kStatError] |= 16; sh.flag |= 4; }
            x.sync();
            break;
        }
        // evaluate: preceding neighbours from the grid, followers from the snapshot
        int nskip = 0;
        for (int it = tid; it < nwork; it += T) {
            const int p = (round == 1) ? it : W.list0[it];
            int i, j, k, a, b;
            face_node<F>(fg, p, i, j, k, a, b);
            // A node whose inputs have not changed since the start of the last
            // full-plane face of this type through its plane would reproduce
            // that evaluation, whose result the grid already holds.
            if (round == 1 && fprev > 0 && max_stamp<F>(J, stamps_of(J, W), i, j, k) < fprev) {
                W.flag[p] = 0;
                nskip++;
                continue;
            }
            Nbhd N;
            load_nbhd<F, true>(J, G, fg, W.O, W.key, p, a, b, i, j, k, N);
            const long long node = (long long)k * nxy + (long long)j * nx + i;
            const float told = W.O[p];
            NodeOut o;
            eval_node<F>(N, G, i, j, k, told, 0.0f, 0.0f, o);
            const float nv = o.stored ? o.tnew : told;
            W.flag[p] = (unsigned char)((o.head ? 1 : 0) | (o.used_in ? 2 : 0) | (o.sets ? 4 : 0) | 8);
            W.tout0[p] = o.tout0;
            W.tout4[p] = o.tout4;
            const float cv = J.time[node];
            unsigned un, uc;
            __builtin_memcpy(&un, &nv, 4);
            __builtin_memcpy(&uc, &cv, 4);
            if (un != uc) {
                J.time[node] = nv;
                if (int *ms = stamps_of(J, W)) ms[node] = fcur;
                W.list2[x.atomic_add(&sh.n[2], 1)] = p;
            }
        }
        if (nskip) x.atomic_add(&sh.st[kStatSkips], nskip);
        if (tid == 0) {
            sh.st[kStatRoundsChange] += 1;
            sh.st[kStatEvals] += nwork;
        }
        if (round == 1) {
            x.sync();
            if (tid == 0) sh.st[kStatTimeChange1] += (int)(x.clock() - tc0);
        }
        x.sync();
        if (tid == 0) sh.n[0] = 0;
        if (qlds)
            for (int w = tid; w < nbits; w += T) qbits[w] = 0u;
        x.sync();
        // queue the followers of every node that changed, once each
        const int nch = sh.n[2];
        for (int it = tid; it < nch; it += T) {
            const int p = W.list2[it];
            const int a = p % fg.wa, b = p / fg.wa;
            const unsigned kp = W.key[p];
            for (int db = -1; db <= 1; db++)
                for (int da = -1; da <= 1; da++) {
                    if (!da && !db) continue;
                    const int aa = a + da, bb = b + db;
                    if (aa < 0 || bb < 0 || aa >= fg.wa || bb >= fg.wb) continue;
                    const int q = bb * fg.wa + aa;
                    if (!key_less(kp, p, W.key[q], q)) continue;
                    bool first;
                    if (qlds) {
                        const unsigned bit = 1u << (q & 31);
                        first = !(x.atomic_or(&qbits[q >> 5], bit) & bit);
                    } else {
                        first = x.atomic_exch(&W.stamp[q], round) != round;
                    }
                    if (first) W.list0[x.atomic_add(&sh.n[0], 1)] = q;
                }
        }
        x.sync();
        if (tid == 0) sh.n[2] = 0;
        x.sync();
    }
    // validity: no node may depend on carried t[0]/t[4]; the last node must set them
    int bad = 0;
    int best = -1;
    unsigned bestk = 0;
    for (int p = tid; p < Fn; p += T) {
        bad |= (W.flag[p] & 2);
        if (best < 0 || key_less(bestk, best, W.key[p], p)) { best = p; bestk = W.key[p]; }
    }
    if (bad) x.atomic_or((unsigned *)&sh.flag, 4u);
    // the last node in face order: tree reduction of (key, index) in LDS
    int *rb = sh.arena;
    unsigned *rk = (unsigned *)(sh.arena + T);
    rb[tid] = best;
    rk[tid] = bestk;
    for (int n = T; n > 1;) {   // any lane count: fold the upper half onto the lower
        const int half = (n + 1) / 2;
        x.sync();
        if (tid < n - half) {
            const int b2 = rb[tid + half];
            const unsigned k2 = rk[tid + half];
            if (b2 >= 0 && (rb[tid] < 0 || key_less(rk[tid], rb[tid], k2, b2))) {
                rb[tid] = b2;
                rk[tid] = k2;
            }
        }
        n = half;
    }
    x.sync();
    if (tid == 0) {
        const int bp = rb[0];
        sh.best = bp;
        if (bp >= 0 && !(W.flag[bp] & 8)) {
            // the last node in face order was skipped: evaluate it once for the
            // t[0]/t[4] it leaves behind (it has no followers, all inputs final)
            int i, j, k, a, b;
            face_node<F>(fg, bp, i, j, k, a, b);
            Nbhd N;
            load_nbhd<F, false>(J, G, fg, W.O, W.key, bp, a, b, i, j, k, N);
            NodeOut o;
            eval_node<F>(N, G, i, j, k, W.O[bp], 0.0f, 0.0f, o);
            // a skipped node must reproduce its stored value and count no head wave
            if ((o.stored && o.tnew != W.O[bp]) || o.head) sh.st[kStatError] |= 32;
            W.flag[bp] = (unsigned char)((o.used_in ? 2 : 0) | (o.sets ? 4 : 0) | 8);
            W.tout0[bp] = o.tout0;
            W.tout4[bp] = o.tout4;
            if (o.used_in) sh.flag |= 4;
        }
        if (bp >= 0 && !(W.flag[bp] & 4)) sh.flag |= 4;
    }
    x.sync();
    if (sh.flag & 4) {
        // restore the face and let LEVEL redo it
        for (int p = tid; p < Fn; p += T) {
            int i, j, k, a, b;
            face_node<F>(fg, p, i, j, k, a, b);
            J.time[(long long)k * nxy + (long long)j * nx + i] = W.O[p];
        }
        if (tid == 0) sh.st[kStatFallbacks] += 1;
        x.sync();
        return false;
    }
    // head waves of the final evaluation of every node
    int hw = 0;
    for (int p = tid; p < Fn; p += T) hw += W.flag[p] & 1;
    if (hw) x.atomic_add(&sh.headw[face_slot(F)], hw);
    if (tid == 0 && sh.best >= 0) { sh.cin0 = W.tout0[sh.best]; sh.cin4 = W.tout4[sh.best]; }
    x.sync();
    return true;
}

template <int F, class X, class SH>
FD_INLINE void sweep_face(X &x, SH &sh, const TableJob &J, const Scratch &W, Geo &G) {
    const int T = x.nthr(), tid = x.tid();
    const int(*B)[2] = G.box;
    FaceGeom fg;
    fg.face = F;
    if (F == TOP || F == BOTTOM) {
        fg.fixed = (F == TOP) ? B[2][0] : B[2][1];
        fg.a0 = B[0][0] + 1; fg.wa = B[0][1] - 1 - fg.a0 + 1;
        fg.b0 = B[1][0] + 1; fg.wb = B[1][1] - 1 - fg.b0 + 1;
    } else if (F == EAST || F == WEST) {
        fg.fixed = (F == EAST) ? B[0][1] : B[0][0];
        fg.a0 = B[1][0] + 1; fg.wa = B[1][1] - 1 - fg.a0 + 1;
        fg.b0 = B[2][0] + 1; fg.wb = B[2][1] - 1 - fg.b0 + 1;
    } else {
        fg.fixed = (F == NORTH) ? B[1][0] : B[1][1];
        fg.a0 = B[0][0] + 1; fg.wa = B[0][1] - 1 - fg.a0 + 1;
        fg.b0 = B[2][0] + 1; fg.wb = B[2][1] - 1 - fg.b0 + 1;
    }
    if (fg.wa <= 0 || fg.wb <= 0) return;
    fg.F = fg.wa * fg.wb;
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    using S = FaceSpec<F>;
    const int di = S::nax == AXX ? S::ndir : 0, dj = S::nax == AXY ? S::ndir : 0, dk = S::nax == AXZ ? S::ndir : 0;
    const long long c0 = x.clock();
    if (tid == 0) { sh.flag = 0; sh.kor = 0u; sh.kand = 0xffffffffu; sh.face_ctr += 1; }
    x.sync();
    const int fcur = sh.face_ctr;
    const int na = FaceSpec<F>::a_ax == AXX ? J.nx : FaceSpec<F>::a_ax == AXY ? J.ny : J.nz;
    const int nb = FaceSpec<F>::b_ax == AXX ? J.nx : FaceSpec<F>::b_ax == AXY ? J.ny : J.nz;
    const bool full = fg.a0 == 0 && fg.wa == na && fg.b0 == 0 && fg.wb == nb;
    int *lastf = lastfull_of(J, W) + F * J.maxdim + fg.fixed;
    const int fprev = (full && J.skip_ok) ? *lastf : 0;
    // gather the pre-face values and the inner-layer keys
    unsigned kor = 0u, kand = 0xffffffffu;
    int fl = 0;
    for (int p = tid; p < fg.F; p += T) {
        int i, j, k, a, b;
        face_node<F>(fg, p, i, j, k, a, b);
        const long long idx = (long long)k * nxy + (long long)j * nx + i;
        const float o = J.time[idx];
        const float kv = J.time[idx + di + dj * nx + dk * nxy];
        W.O[p] = o;
        const unsigned kk = order_key(kv);
        W.key[p] = kk;
        W.val[p] = p;
        kor |= kk;
        kand &= kk;
        if (!((double)o < kBig)) fl |= 1;
        if (kv != kv) fl |= 2;
    }
    x.atomic_or(&sh.kor, kor);
    x.atomic_and(&sh.kand, kand);
    if (fl) x.atomic_or((unsigned *)&sh.flag, (unsigned)fl);
    x.sync();
    if (sh.flag & 2) {
        if (tid == 0) sh.st[kStatError] |= 1;   // NaN keys: the reference's qsort order is undefined
    }
    const bool try_change = !(sh.flag & 1);
    x.sync();
    const long long c1 = x.clock();
    if (tid == 0) sh.st[kStatTimeGather] += (int)(c1 - c0);
    bool done = false;
    if (try_change) {
        done = face_change<F>(x, sh, J, W, G, fg, fcur, fprev);
        if (done && tid == 0) sh.st[kStatFacesChange] += 1;
    }
    const long long c2 = x.clock();
    if (tid == 0) sh.st[kStatTimeChange] += (int)(c2 - c1);
    if (!done) {
        // LEVEL needs ranks only; the sort may leave the keys permuted
        const bool sw = radix_sort(x, sh, W.key, W.val, W.key2, W.val2, fg.F);
        const int *ord = sw ? W.val2 : W.val;
        const long long c3 = x.clock();
        face_level<F>(x, sh, J, W, G, fg, ord, fcur);
        if (tid == 0) {
            sh.st[kStatFacesLevel] += 1;
            sh.st[kStatTimeSort] += (int)(c3 - c2);
            sh.st[kStatTimeLevel] += (int)(x.clock() - c3);
        }
    }
    // Only a CHANGE face is a valid reference for later skips: there every node
    // started below 1e9 (no new-edge or new-face trials) and set its own
    // t[0]/t[4]; a LEVEL face's results also depend on the carried registers.
    if (full && done && tid == 0) *lastf = fcur;
    x.sync();
}


}  // namespace fdsweep

#include "sweep_wave.h"

namespace fdsweep {

// ------------------------------------------------------------- one table

// Solves table J on one workgroup, or as the leader of a team of K
// workgroups (tm: the team's block) whose other members run team_helper; the
// caller releases the team when this returns.
template <class X, class SH>
FD_HD void solve_table(X &x, SH &sh, const TableJob &J, const Scratch &W, TeamCtl *tm = nullptr, int K = 1) {
    const int T = x.nthr(), tid = x.tid();
    const long long N = (long long)J.nx * J.ny * J.nz;
    if (int *ms = stamps_of(J, W))
        for (long long q = tid; q < N; q += T) ms[q] = 0;
    int *lf = lastfull_of(J, W);
    for (int q = tid; q < 6 * J.maxdim; q += T) lf[q] = 0;
    if (J.init_mode != 1) {
        for (long long q = tid; q < N; q += T) J.time[q] = 1.0e10f;
        x.sync();
        const float s000 = J.slow[J.src_sidx];
        for (int q = tid; q < J.ninit; q += T)
            J.time[J.init_idx[q]] = J.init_mode == 0 ? (float)((double)s000 * J.init_dist[q]) : J.init_val[q];
    }
    if (J.geo_fill) {
        const GeoParams gp{J.ny, J.nz, J.h, J.dq, J.df, J.z0r, J.sinq};
        const long long ne = geo_items(J.ny, J.nz);
        for (long long e = tid; e < ne; e += T) geo_fill_item(gp, e, J.geo_fill);
        x.sync();
    }
    // the row and layer tables into LDS
    const long long nrl = geo_rows_doubles(J.ny) + geo_layers_doubles(J.nz);
    if (nrl > SH::kGeoLds) {   // never sent by the host (it picks a build that holds them)
        if (tid == 0) {
            for (int s = 0; s < kStatCount; s++) sh.st[s] = 0;
            sh.st[kStatError] = 64;
        }
        x.sync();
        for (int q = tid; q < kStatCount; q += T) J.stats[q] = sh.st[q];
        x.sync();
        return;
    }
    Geo G;
    load_geo(x, sh, J, G);
    if (tid == 0) {
        for (int s = 0; s < kStatCount; s++) sh.st[s] = 0;
        sh.cin0 = 0.0f;
        sh.cin4 = 0.0f;
        sh.face_ctr = 0;
        sh.st[kStatTeam] = tm ? K : 1;
    }
    x.sync();
    int x1 = J.box[0][0], x2 = J.box[0][1], y1 = J.box[1][0], y2 = J.box[1][1], z1 = J.box[2][0], z2 = J.box[2][1];
    int dx1 = J.dflag[0], dx2 = J.dflag[1], dy1 = J.dflag[2], dy2 = J.dflag[3], dz1 = J.dflag[4], dz2 = J.dflag[5];
    int rad0 = 1, radius = J.radius0, srcwall = J.srcwall, reverse = J.reverse;
    int gen = 0;   // team barriers passed
    const int nx = J.nx, ny = J.ny, nz = J.nz;
    int passes = 0;
#define FD_SETBOX()                                  \
    do {                                             \
        G.box[0][0] = x1; G.box[0][1] = x2;          \
        G.box[1][0] = y1; G.box[1][1] = y2;          \
        G.box[2][0] = z1; G.box[2][1] = z2;          \
    } while (0)
    while (reverse > -1) {
        if (tid == 0)
            for (int w = 0; w < 7; w++) sh.headw[w] = 0;
        x.sync();
        passes++;
        const bool wave = (J.wave & (passes == 1 ? 1 : 2)) && W.wst && N <= W.wn && N < (1LL << 29) &&
                          wave_fits(nx, ny, nz, SH::kArena);
        if (wave) {
            const int b0[3][2] = {{x1, x2}, {y1, y2}, {z1, z2}};
            const int f0[6] = {dx1, dx2, dy1, dy2, dz1, dz2};
            // the whole pass at once; on failure the host solves the table face by face
            if (!wave_pass(x, sh, J, W, G, b0, f0, rad0, radius, tm, tm ? K : 1, 0, &gen)) break;
            dx1 = dx2 = dy1 = dy2 = dz1 = dz2 = 0;
        }
        // no wavefront pass now, none later: the team's other members leave
        if (tm && (!wave || !(J.wave & 2))) team_release(x, tm);
        while (rad0 && (dx1 || dx2 || dy1 || dy2 || dz1 || dz2)) {
            for (int g = 1; g <= J.grow[0]; g++)
                if (dz1) { FD_SETBOX(); sweep_face<TOP>(x, sh, J, W, G); if (z1 == 0) dz1 = 0; z1--; }
            for (int g = 1; g <= J.grow[1]; g++)
                if (dz2) { FD_SETBOX(); sweep_face<BOTTOM>(x, sh, J, W, G); if (z2 == nz - 1) dz2 = 0; z2++; }
            for (int g = 1; g <= J.grow[2]; g++)
                if (dx2) { FD_SETBOX(); sweep_face<EAST>(x, sh, J, W, G); if (x2 == nx - 1) dx2 = 0; x2++; }
            for (int g = 1; g <= J.grow[3]; g++)
                if (dx1) { FD_SETBOX(); sweep_face<WEST>(x, sh, J, W, G); if (x1 == 0) dx1 = 0; x1--; }
            for (int g = 1; g <= J.grow[4]; g++)
                if (dy1) { FD_SETBOX(); sweep_face<NORTH>(x, sh, J, W, G); if (y1 == 0) dy1 = 0; y1--; }
            for (int g = 1; g <= J.grow[5]; g++)
                if (dy2) { FD_SETBOX(); sweep_face<SOUTH>(x, sh, J, W, G); if (y2 == ny - 1) dy2 = 0; y2++; }
            radius++;
            if (radius == J.maxrad) rad0 = 0;
        }
        x.sync();
        int hw[7];
        for (int w = 0; w < 7; w++) hw[w] = sh.headw[w];
        if (passes == 1 && tid == 0)
            for (int w = 0; w < 7; w++) sh.st[kStatHeadw0 + w] = hw[w];
        // sphfd.c:5570-5676
        if (hw[1] == 0 && hw[2] == 0 && hw[3] == 0 && hw[4] == 0 && hw[5] == 0 && hw[6] == 0) {
            reverse = 0;
        } else {
            int head = 0;
            for (int w = 1; w <= 6; w++)
                if (hw[w] > 0 && hw[w] > head) { head = hw[w]; srcwall = w; }
            if (J.headpref > 0 && hw[J.headpref] > 0) srcwall = J.headpref;
            dx1 = dx2 = dy1 = dy2 = dz1 = dz2 = 1;
            rad0 = 1;
            radius = 1;
            if (srcwall == 1) x2 = 1; else { x2 = nx; dx2 = 0; }
            if (srcwall == 2) x1 = nx - 2; else { x1 = -1; dx1 = 0; }
            if (srcwall == 3) y2 = 1; else { y2 = ny; dy2 = 0; }
            if (srcwall == 4) y1 = ny - 2; else { y1 = -1; dy1 = 0; }
            if (srcwall == 5) z2 = 1; else { z2 = nz; dz2 = 0; }
            if (srcwall == 6) z1 = nz - 2; else { z1 = -1; dz1 = 0; }
            if (tid == 0) {
                sh.st[kStatLastWall] = srcwall;
                if (reverse > 0) sh.st[kStatRestarts] += 1;
            }
        }
        reverse--;
        x.sync();
    }
#undef FD_SETBOX
    if (tid == 0) sh.st[kStatPasses] = passes;
    x.sync();
    for (int q = tid; q < kStatCount; q += T) J.stats[q] = sh.st[q];
    x.sync();
}

}  // namespace fdsweep
