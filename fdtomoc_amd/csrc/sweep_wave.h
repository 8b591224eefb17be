// sweep_wave.h -- one pass of the expanding-box sweep (the first pass or a
// reverse pass of sphfd.c:1008-5678) run as a dataflow wavefront over the
// whole table instead of face by face.  Included by sweep_solver.h.
//
// The reference visits the faces of a pass in a fixed order and the nodes of a
// face in ascending (inner-layer time, fill index) order; a node reads the
// current values of its 3x3 in-face and 3x3 inner-layer neighbours.  So every
// dependency is between two nodes of one 3x3x3 cube, and the sequential result
// is reproduced by any order in which a node is evaluated once every
// neighbour that precedes it is final and before any neighbour that follows it
// is touched.  u precedes v when u's face comes first, or when both lie on one
// face and u's (key, fill index) is lower.
//
// Face membership is a closed form.  Replaying the controller's box growth
// (the same loop as solve_table) gives, for every plane of each axis, the face
// event that sweeps it, -1 for planes inside the starting box and kWaveNever
// for planes the pass never reaches.  A node belongs to the event with the
// largest index among its three planes' events: at that event the other two
// coordinates are already inside the box, and at the earlier ones they are
// not.  The order key of a face node is its inner neighbour's time; the inner
// neighbours of a node and of its in-face neighbours lie in the node's cube on
// an earlier face, so they are final before the node can be found ready.
//
// Rounds (bulk-synchronous).  Every node of the pass has a counter word: its
// earlier-face predecessors still to finalise (set at the start of the pass
// from the event tables) and, beside it, how many of its in-face predecessors
// have finalised.  A node evaluated in round R (exactly as a LEVEL face
// evaluates it, sweep_core.h eval_node) updates the counters of its
// neighbours that follow it; the update that leaves no earlier-face
// predecessor and all in-face predecessors final queues the node for round
// R + 1.  The number of in-face predecessors needs the order keys, which are
// final by then.  Two neighbours are never ready in the same round (one of
// them precedes the other), so a round reads only values final before it,
// and its nodes can be evaluated in any order by any number of workgroups.
//
// Teams.  A table's passes run on one workgroup, or on a team of K
// workgroups (on K CUs) that share each round's list: member m evaluates the
// m-th K-th of it.  The list sizes, the requests below and the per-face key
// counters then live in the team's block in device memory (TeamCtl), and the
// members meet at a barrier in device memory (team_sync: agent-scope release
// before arriving, acquire after leaving) twice per round: after the leader
// has swapped the lists, and after every member has evaluated its share.
// The leader (member 0) owns the table: it runs the controller and every
// face-by-face pass, opens each wavefront pass for the others, and resolves
// the carried-register requests between rounds.
//
// Carried registers: fdsph2d also tests t[0]/t[4] as left by whichever node
// ran before it in the sequential order (sphfd.c:6116-6121).  A node that
// needs them (needs_carry) waits in a small request list until every key of
// its face is final (per-face counters of keys not yet final), then a scan of
// the face finds the node with the largest (key, fill) below its own: if that
// node is final and set the registers, its t[0]/t[4] are the carry; if it did
// not set them it passed its own carry on, so the search continues below it;
// below the first node of a face it continues at the last node of the
// previous face, and before the first face it is the pass's carry-in.
#pragma once

namespace fdsweep {

constexpr int kWaveNever = 0x7fff0000;   // plane not swept in this pass (above every event)
constexpr int kWaveCarryCap = 64;        // carried-register requests waiting at once
constexpr int kWaveEvInts = 6;           // per event: 3 packed box pairs, type, face size, keys pending
constexpr int kWaveCarryInts = 8;        // per request: node, event, bound (2), candidate, state, cand. event, -
constexpr unsigned kWaveRoundLimit = 1u << 29;   // a slot's rounds are renumbered from 1 past this
constexpr int kWaveLocalMin = 1024;      // arena words left for grouping a round's nodes by face type
constexpr int kTeamMaxEv = 3712;         // events of a pass in a team's block (at most nx + ny + nz)
constexpr int kTeamMax = 16;             // workgroups in a team
// a member waiting longer than this (X::clock ticks of 10 ns) gives the pass
// up, unless the team's block sets another limit
constexpr long long kTeamWaitTicks = 2000000000LL;

// Arena words the wavefront needs for a grid (planes' events, events, requests).
FD_INLINE long long wave_arena_words(int nx, int ny, int nz) {
    return (long long)(nx + ny + nz) * (1 + kWaveEvInts) + (long long)kWaveCarryCap * kWaveCarryInts;
}

// Whether a grid's passes can run as wavefronts in an arena of `arena` words
// (the event tables plus room to group a round's nodes).
FD_INLINE bool wave_fits(int nx, int ny, int nz, int arena) {
    return wave_arena_words(nx, ny, nz) + kWaveLocalMin <= arena && nx + ny + nz <= kTeamMaxEv;
}

// Control words in SharedT::wc (the list sizes kWcNCur/kWcNNext and
// kWcNClose/kWcNClose1, indexed by round parity, kWcErr and kWcWaiting are
// the team's words in TeamCtl::wc when the pass runs on a team).
enum {
    kWcNev = 0, kWcNCur, kWcNNext, kWcDone, kWcTotal, kWcErr, kWcScanBad, kWcWaiting, kWcTeamOk,
    kWcNClose, kWcNClose1,
    kWcTypeCnt = 16,   // 6 counters of the nodes being grouped
    kWcTypeOff = 22,   // 6 offsets into the grouped nodes
};

// A team's block in device memory, one per table solved by a team.  The
// words up to `W` are cleared before every launch; the leader writes the rest
// when it opens a pass.
//
// Members on different dies reach it through different L2 caches.  A word
// that a member updates with an atomic while another writes plain stores
// into the same cache line is not safe from the latter's write-back, so every
// word that members change while others run (the barrier and pass words, the
// list sizes, the key counters, the requests, the counts) is only ever
// written and read with atomics, and the one part written with plain stores
// (the pass parameters, by the leader before it opens a pass) has cache
// lines of its own.
struct TeamCtl {
    alignas(128) int seq;   // wavefront passes the leader has opened
    int release;      // 1: the leader opens no more passes (the others leave)
    int abort;        // 1: a member waited too long (every member gives the pass up)
    int bar_count;    // members arrived at the barrier
    int bar_gen;      // barriers completed
    int inpass;       // members other than the leader inside a pass
    long long wait_ticks;   // how long a member waits before it gives up (X::clock ticks)
    alignas(128) Scratch W;   // the leader's slot: the pass's node state and lists
    int box0[3][2], flags0[6], rad0, radius, stamp0;
    unsigned base;    // the pass's first round number
    alignas(128) int wc[16];  // list sizes, waiting requests, error
    int headw[8];     // head waves counted by the other members
    int done, seeds, carry;   // evaluations (all members), the others' seeds and requests
    alignas(128) int pk[kTeamMaxEv];   // per event: keys of its face not yet final
    alignas(128) int car[kWaveCarryCap * kWaveCarryInts];
};

// The list a round fills: nodes queued for the next round, with its size
// word.  In a team the entries are written with atomic stores (they bypass
// the caches, so the barrier after the closing phase needs no write-back).
struct WaveQ {
    int *next;
    int *cnt;
    long long cap;
    bool team;
};

// Barrier of a team's K members: every store a member made before it is
// visible to every member after it.  false: a member gave up waiting (the
// pass is abandoned by all).
//
// One lane per workgroup does the cache maintenance: every wave first waits
// for its own stores to reach the CU's L2 (X::wait_stores), the workgroup
// barrier orders that before lane 0's release (which writes the L2 back for
// the other dies), and lane 0's acquire (which invalidates the CU's L1 and
// the lines of other dies' memory in its L2) comes before the barrier that
// lets the workgroup read.  The spin reads are relaxed: an acquire per
// iteration would invalidate the caches on every poll.
template <class X>
FD_INLINE void team_publish(X &x) {
    x.wait_stores();
    x.sync();
    if (x.tid() == 0) x.fence_release();
}

//
// light: the words written since the last barrier that members read after
// this one were all written and are all read with atomics, so it only
// orders (each wave waits for its stores) and neither writes the L2 back nor
// invalidates it.
template <class X>
FD_INLINE bool team_sync(X &x, TeamCtl *tm, int K, int &gen, int *okw, bool light = false) {
    if (light) {
        x.wait_stores();
        x.sync();
    } else {
        team_publish(x);
    }
    if (x.tid() == 0) {
        int ok = 1;
        if (x.load_relaxed(&tm->abort)) {
            ok = 0;
        } else if (x.atomic_add(&tm->bar_count, 1) == K - 1) {
            // the last to arrive: it acquires the others' stores, then
            // releases them with its own to everyone who sees the new count
            if (!light) x.fence_acquire();
            x.atomic_exch(&tm->bar_count, 0);
            if (light) x.store_relaxed(&tm->bar_gen, gen + 1);
            else x.store_release(&tm->bar_gen, gen + 1);
        } else {
            const long long t0 = x.clock(), lim = tm->wait_ticks > 0 ? tm->wait_ticks : kTeamWaitTicks;
            while (x.load_relaxed(&tm->bar_gen) == gen) {
                if (x.load_relaxed(&tm->abort)) { ok = 0; break; }
                if (x.clock() - t0 > lim) { x.store_release(&tm->abort, 1); ok = 0; break; }
                x.pause();
            }
            if (!light) x.fence_acquire();
        }
        *okw = ok;
    }
    gen++;
    x.sync();
    const int ok = *okw;
    x.sync();
    return ok != 0;
}

// The leader opens no more passes: the other members leave the table.
template <class X>
FD_INLINE void team_release(X &x, TeamCtl *tm) {
    team_publish(x);
    if (x.tid() == 0) x.store_release(&tm->release, 1);
    x.sync();
}

// status word: bit0 final, bit1 set t[0]/t[4], bits 2.. the round
FD_INLINE bool wv_final_in(unsigned s, unsigned base) { return (s & 1u) && (s >> 2) >= base; }
FD_INLINE bool wv_final_before(unsigned s, unsigned base, unsigned R) {
    return (s & 1u) && (s >> 2) >= base && (s >> 2) < R;
}

// Views of the pass's event tables in the workgroup arena (the key counters
// and requests in the team's block when a team runs the pass).
struct WaveTabs {
    int *mx, *my, *mz;   // plane -> event
    int *ev;             // kWaveEvInts per event
    int *car;            // kWaveCarryInts per request
    int *pkb;            // keys not yet final of event e at pkb[pks * e]
    int pks;
    FD_INLINE int type(int e) const { return ev[kWaveEvInts * e + 3]; }
    FD_INLINE int fsize(int e) const { return ev[kWaveEvInts * e + 4]; }
    FD_INLINE int *pk(int e) const { return pkb + pks * e; }
    FD_INLINE void box(int e, int B[3][2]) const {
        for (int a = 0; a < 3; a++) {
            const unsigned w = (unsigned)ev[kWaveEvInts * e + a];
            B[a][0] = (int)(w & 0xffffu) - 1;
            B[a][1] = (int)(w >> 16) - 1;
        }
    }
};

FD_INLINE int imax3(int a, int b, int c) {
    const int m = a > b ? a : b;
    return m > c ? m : c;
}

// Face geometry of event e (as sweep_face computes it from the box).
FD_INLINE void wave_face_geom(const WaveTabs &WT, int e, FaceGeom &fg) {
    int B[3][2];
    WT.box(e, B);
    const int F = WT.type(e);
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
    fg.F = (fg.wa > 0 && fg.wb > 0) ? fg.wa * fg.wb : 0;
}

// Normal axis and inward direction of a face type (FaceSpec<F>::nax / ndir).
FD_INLINE int face_nax(int F) { return (F == TOP || F == BOTTOM) ? 2 : (F == EAST || F == WEST) ? 0 : 1; }
FD_INLINE int face_ndir(int F) { return (F == TOP || F == WEST || F == NORTH) ? 1 : -1; }

// Coordinates of fill index p of a face, and the node's fill index.
FD_INLINE void face_coords(const FaceGeom &fg, int p, int &i, int &j, int &k) {
    const int a = p % fg.wa, b = p / fg.wa;
    const int A = fg.a0 + a, B = fg.b0 + b;
    if (fg.face == TOP || fg.face == BOTTOM) { i = A; j = B; k = fg.fixed; }
    else if (fg.face == EAST || fg.face == WEST) { i = fg.fixed; j = A; k = B; }
    else { i = A; j = fg.fixed; k = B; }
}
FD_INLINE int face_fill(const FaceGeom &fg, int i, int j, int k) {
    if (fg.face == TOP || fg.face == BOTTOM) return (j - fg.b0) * fg.wa + (i - fg.a0);
    if (fg.face == EAST || fg.face == WEST) return (k - fg.b0) * fg.wa + (j - fg.a0);
    return (k - fg.b0) * fg.wa + (i - fg.a0);
}

// Appends ready node v (face type F) to the next round's list (a list over
// its capacity fails the pass when the round starts).
template <class X>
FD_INLINE void wave_queue(X &x, const WaveQ &Q, long long v, int F) {
    const int pos = x.atomic_add(Q.cnt, 1);
    const int w = (int)((unsigned)v | ((unsigned)F << 29));
    if (pos < Q.cap) {
        if (Q.team) x.store_relaxed(&Q.next[pos], w);
        else Q.next[pos] = w;
    }
}

// An entry of a list (read with an atomic load in a team, see WaveQ).
template <class X>
FD_INLINE int wave_entry(X &x, const int *list, long long i, bool team) {
    return team ? x.load_relaxed(list + i) : list[i];
}

// Node index -> coordinates (indices below 2^29).
FD_INLINE void node_ijk(const TableJob &J, long long v, int &i, int &j, int &k) {
    const unsigned nx = (unsigned)J.nx, nxy = (unsigned)J.nx * (unsigned)J.ny, u = (unsigned)v;
    k = (int)(u / nxy);
    const unsigned r = u - (unsigned)k * nxy;
    j = (int)(r / nx);
    i = (int)(r - (unsigned)j * nx);
}

// Events of the planes through a node's 3x3x3 cube (-2: outside the grid).
FD_INLINE void cube_events(const TableJob &J, const WaveTabs &WT, int i, int j, int k, int ex[3], int ey[3],
                           int ez[3]) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int ci = i + d - 1, cj = j + d - 1, ck = k + d - 1;
        ex[d] = (ci >= 0 && ci < J.nx) ? WT.mx[ci] : -2;
        ey[d] = (cj >= 0 && cj < J.ny) ? WT.my[cj] : -2;
        ez[d] = (ck >= 0 && ck < J.nz) ? WT.mz[ck] : -2;
    }
}

// Earlier-face predecessors of a node of event e: its neighbours swept by an
// earlier event of the pass.
FD_INLINE int earlier_preds(const int ex[3], const int ey[3], const int ez[3], int e) {
    int n = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) {
        if (q == 13) continue;
        const int dx = q % 3, dy = (q / 3) % 3, dz = q / 9;
        if (ex[dx] == -2 || ey[dy] == -2 || ez[dz] == -2) continue;
        const int eu = imax3(ex[dx], ey[dy], ez[dz]);
        n += (eu >= 0 && eu < e) ? 1 : 0;
    }
    return n;
}

// In-face predecessors of node (i, j, k) of event e: in-plane neighbours on
// the same face with a lower (key, fill).  Needs the keys final: the node's
// earlier-face predecessors include every key source read here.
FD_INLINE int inface_preds(const TableJob &J, const WaveTabs &WT, int i, int j, int k, int e, const int ex[3],
                           const int ey[3], const int ez[3]) {
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    FaceGeom fg;
    wave_face_geom(WT, e, fg);
    const int F = fg.face, nax = face_nax(F), nd = face_ndir(F);
    const int aax = (F == EAST || F == WEST) ? 1 : 0, bax = (F == TOP || F == BOTTOM) ? 1 : 2;
    const int n3[3] = {J.nx, J.ny, J.nz};
    float kt[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
        int c[3] = {i, j, k};
        c[aax] += q % 3 - 1;
        c[bax] += q / 3 - 1;
        c[nax] += nd;
#pragma unroll
        for (int a = 0; a < 3; a++) c[a] = c[a] < 0 ? 0 : (c[a] >= n3[a] ? n3[a] - 1 : c[a]);
        kt[q] = J.time[(long long)c[2] * nxy + (long long)c[1] * nx + c[0]];
    }
    const unsigned long long kv = ((unsigned long long)order_key(kt[4]) << 32) | (unsigned)face_fill(fg, i, j, k);
    int n = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if (q == 4) continue;
        int c[3] = {i, j, k};
        c[aax] += q % 3 - 1;
        c[bax] += q / 3 - 1;
        if (c[aax] < 0 || c[aax] >= n3[aax] || c[bax] < 0 || c[bax] >= n3[bax]) continue;
        const int d3[3] = {c[0] - i + 1, c[1] - j + 1, c[2] - k + 1};
        if (imax3(ex[d3[0]], ey[d3[1]], ez[d3[2]]) != e) continue;
        const unsigned long long ku = ((unsigned long long)order_key(kt[q]) << 32) |
                                      (unsigned)face_fill(fg, c[0], c[1], c[2]);
        n += ku < kv ? 1 : 0;
    }
    return n;
}

// Collective: the largest (key << 32 | fill) below `bound` over the nodes of
// event e (setters only if `setters`), or ~0ull if there is none.  Also
// checks that every key it reads is final; if one is not, sets kWcScanBad.
template <class X, class SH>
FD_NOINLINE unsigned long long wave_scan(X &x, SH &sh, const TableJob &J, const Scratch &W, const WaveTabs &WT,
                                       int e, unsigned long long bound, bool setters, unsigned base) {
    const int T = x.nthr(), tid = x.tid();
    FaceGeom fg;
    wave_face_geom(WT, e, fg);
    const int F = fg.face, nax = face_nax(F), nd = face_ndir(F);
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    if (tid == 0) { sh.wbest = 0ull; sh.wc[kWcScanBad] = 0; }
    x.sync();
    unsigned long long best = ~0ull;
    bool bad = false;
    for (int p = tid; p < fg.F; p += T) {
        int i, j, k;
        face_coords(fg, p, i, j, k);
        int ii = i, jj = j, kk = k;
        if (nax == 0) ii += nd; else if (nax == 1) jj += nd; else kk += nd;
        const long long w = (long long)kk * nxy + (long long)jj * nx + ii;
        const int ew = imax3(WT.mx[ii], WT.my[jj], WT.mz[kk]);
        if (ew >= 0 && ew < kWaveNever && !wv_final_in(W.wst[w], base)) bad = true;
        const unsigned long long key = ((unsigned long long)order_key(J.time[w]) << 32) | (unsigned)p;
        if (key >= bound) continue;
        if (setters) {
            const unsigned s = W.wst[(long long)k * nxy + (long long)j * nx + i];
            if (!(wv_final_in(s, base) && (s & 2u))) continue;
        }
        if (best == ~0ull || key > best) best = key;
    }
    if (best != ~0ull) x.atomic_max_u64(&sh.wbest, best + 1);   // +1: 0 stays "none"
    if (bad) sh.wc[kWcScanBad] = 1;
    x.sync();
    const unsigned long long r = sh.wbest;
    x.sync();
    return (r == ~0ull || r == 0) ? ~0ull : r - 1;
}

// Node u's earlier-face predecessors are all final: count its in-face
// predecessors (their keys are final now), record the count, and queue u if
// they are all final already.
template <class X>
FD_INLINE void wave_close(X &x, const TableJob &J, const Scratch &W, const WaveTabs &WT, const WaveQ &Q, long long u) {
    int ui, uj, uk;
    node_ijk(J, u, ui, uj, uk);
    int fx[3], fy[3], fz[3];
    cube_events(J, WT, ui, uj, uk, fx, fy, fz);
    const int eu = imax3(fx[1], fy[1], fz[1]);
    const int p2 = inface_preds(J, WT, ui, uj, uk, eu, fx, fy, fz);
    const int o2 = x.atomic_add(&W.wcnt[u], (p2 << 16) | (1 << 24));
    if (((o2 >> 8) & 255) == p2) wave_queue(x, Q, u, WT.type(eu));
}

// Evaluates ready node v of face type F as a LEVEL face would, then updates
// the counters of its neighbours that follow it and queues those it leaves
// ready.
//
// A node whose earlier-face predecessors this evaluation closes needs its
// in-face predecessors counted from its inner-layer keys.  One workgroup
// does that at once: the keys written this round by its other waves passed
// through the same CU and are visible.  In a team (CL set) a key may have
// been written this round on another CU, visible only after the next
// barrier, so the node goes to the closing list CL instead and is counted
// in the round's closing phase (wave_close).
template <int F, class X, class SH>
FD_INLINE void wave_eval(X &x, SH &sh, const TableJob &J, const Scratch &W, Geo &G, const WaveTabs &WT,
                         const WaveQ &Q, const WaveQ *CL, long long v, unsigned base, unsigned R, int stamp0) {
    const long long nx = J.nx, nxy = (long long)J.nx * J.ny;
    int i, j, k;
    node_ijk(J, v, i, j, k);
    int ex[3], ey[3], ez[3];
    cube_events(J, WT, i, j, k, ex, ey, ez);
    const int e = imax3(ex[1], ey[1], ez[1]);
    WT.box(e, G.box);
    FaceGeom fg{};
    Nbhd N;
    load_nbhd<F, false>(J, G, fg, nullptr, nullptr, 0, 0, 0, i, j, k, N);
    const float told = N.t[4];
    float tin0 = 0.0f, tin4 = 0.0f;
    if (needs_carry<F>(N, G, i, j, k, told)) {
        // its resolved request, if any
        int slot = -1;
        for (int c = 0; c < kWaveCarryCap; c++) {
            const int *q = WT.car + kWaveCarryInts * c;
            if (x.load_relaxed(&q[5]) == 2 && x.load_relaxed(&q[0]) == (int)v) { slot = c; break; }
        }
        if (slot >= 0) {
            // the resolved carry-in: the leader stored it atomically
            tin0 = __builtin_bit_cast(float, x.load_relaxed((const int *)&W.wt0[v]));
            tin4 = __builtin_bit_cast(float, x.load_relaxed((const int *)&W.wt4[v]));
            x.atomic_exch(&WT.car[kWaveCarryInts * slot + 5], 0);   // free the request
        } else {
            // file a request (the node is queued again when it is resolved)
            FaceGeom g2;
            wave_face_geom(WT, e, g2);
            const int nax = face_nax(F), nd = face_ndir(F);
            int ii = i, jj = j, kk = k;
            if (nax == 0) ii += nd; else if (nax == 1) jj += nd; else kk += nd;
            const unsigned key = order_key(J.time[(long long)kk * nxy + (long long)jj * nx + ii]);
            int got = -1;
            for (int c = 0; c < kWaveCarryCap && got < 0; c++)
                if (x.atomic_cas((unsigned *)&WT.car[kWaveCarryInts * c + 5], 0u, 3u) == 0u) got = c;
            if (got < 0) {   // no free request: try again next round
                wave_queue(x, Q, v, F);
                return;
            }
            int *q = WT.car + kWaveCarryInts * got;
            x.atomic_exch(&q[0], (int)v);
            x.atomic_exch(&q[1], e);
            x.atomic_exch(&q[2], face_fill(g2, i, j, k));
            x.atomic_exch(&q[3], (int)key);
            x.atomic_exch(&q[4], -1);
            x.atomic_exch(&q[6], e);
            x.atomic_add(&sh.st[kStatWaveCarry], 1);
            x.atomic_exch(&q[5], 1);   // waiting
            return;
        }
    }
    NodeOut o;
    eval_node<F>(N, G, i, j, k, told, tin0, tin4, o);
    if (o.stored) {
        J.time[v] = o.tnew;
        if (int *ms = stamps_of(J, W)) ms[v] = stamp0 + e;
        if (o.head) x.atomic_add(&sh.headw[face_slot(F)], 1);
    }
    W.wt0[v] = o.tout0;
    W.wt4[v] = o.tout4;
    W.wst[v] = (R << 2) | (o.sets ? 2u : 0u) | 1u;
    x.atomic_add(&sh.wc[kWcDone], 1);
    // keys this node provides: the face nodes whose inner neighbour it is
#pragma unroll
    for (int d = 0; d < 6; d++) {
        const int ax = d >> 1, dd = (d & 1) ? 2 : 0;   // the axis neighbour at offset dd - 1
        const int eu = ax == 0 ? imax3(ex[dd], ey[1], ez[1]) : ax == 1 ? imax3(ex[1], ey[dd], ez[1])
                                                                     : imax3(ex[1], ey[1], ez[dd]);
        const bool in = ax == 0 ? ex[dd] != -2 : ax == 1 ? ey[dd] != -2 : ez[dd] != -2;
        if (!in || eu < 0 || eu >= kWaveNever) continue;
        const int Fu = WT.type(eu);
        if (face_nax(Fu) == ax && face_ndir(Fu) == 1 - dd) x.atomic_add(WT.pk(eu), -1);
    }
    // the neighbours that follow it: those of later faces, and those of this
    // face with a higher (key, fill) -- their keys are the inner-layer times
    // already loaded (N.t[9..17])
    FaceGeom fgv;
    wave_face_geom(WT, e, fgv);
    const int fv = face_fill(fgv, i, j, k);
    const unsigned long long kv = ((unsigned long long)order_key(N.t[13]) << 32) | (unsigned)fv;
    using S = FaceSpec<F>;
    unsigned want = 0, same = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) {
        if (q == 13) continue;
        const int dx = q % 3, dy = (q / 3) % 3, dz = q / 9;
        if (ex[dx] == -2 || ey[dy] == -2 || ez[dz] == -2) continue;
        const int eu = imax3(ex[dx], ey[dy], ez[dz]);
        if (eu < e || eu >= kWaveNever) continue;
        if (eu > e) { want |= 1u << q; continue; }
        // same face: an in-plane neighbour
        const int d3[3] = {dx - 1, dy - 1, dz - 1};
        const int da = d3[S::a_ax], db = d3[S::b_ax];
        const unsigned long long ku = ((unsigned long long)order_key(N.t[9 + (db + 1) * 3 + (da + 1)]) << 32) |
                                      (unsigned)(fv + db * fgv.wa + da);
        if (ku > kv) { want |= 1u << q; same |= 1u << q; }
    }
    // counter word: bits 0-7 earlier-face predecessors still open, 8-15 in-face
    // predecessors final, 16-23 the in-face predecessor count, bit 24 set once
    // that count is known (by the update that closed the earlier-face ones)
    int old[27];
#pragma unroll
    for (int q = 0; q < 27; q++) {
        old[q] = 0;
        if (!(want & (1u << q))) continue;
        const int dx = q % 3, dy = (q / 3) % 3, dz = q / 9;
        old[q] = x.atomic_add(&W.wcnt[(long long)(k + dz - 1) * nxy + (long long)(j + dy - 1) * nx + (i + dx - 1)],
                              (same & (1u << q)) ? 256 : -1);
    }
    unsigned closed = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) {
        if (!(want & (1u << q))) continue;
        const int o = old[q];
        const int dx = q % 3, dy = (q / 3) % 3, dz = q / 9;
        if (same & (1u << q)) {
            // ready when the count is known and this was the last in-face predecessor
            if (((o >> 24) & 1) && ((o >> 8) & 255) + 1 == ((o >> 16) & 255))
                wave_queue(x, Q, (long long)(k + dz - 1) * nxy + (long long)(j + dy - 1) * nx + (i + dx - 1),
                           WT.type(e));
        } else if ((o & 255) == 1) {
            closed |= 1u << q;   // this update closed its earlier-face predecessors
        }
    }
    // for those: count their in-face predecessors (now, or in the closing phase)
    while (closed) {
        const int q = __builtin_ctz(closed);
        closed &= closed - 1;
        const int ui = i + q % 3 - 1, uj = j + (q / 3) % 3 - 1, uk = k + q / 9 - 1;
        const long long u = (long long)uk * nxy + (long long)uj * nx + ui;
        if (CL) wave_queue(x, *CL, u, 0);
        else wave_close(x, J, W, WT, Q, u);
    }
}

// Runs one pass from the controller state (box, growth flags, rad0, radius)
// as a wavefront, alone (tm null) or as member `member` of a team of K
// workgroups.  The leader (member 0, also the lone workgroup) owns the pass;
// the others are called by team_helper with the leader's slot W and the
// pass's parameters.  Returns false if the pass could not be completed (list
// capacity, a request that never resolves, a member that waited too long:
// the host then solves the table face by face).  Every member reaches the
// same number of team barriers: each decision to leave the rounds is taken on
// values every member reads after the same barrier.
template <class X, class SH>
FD_INLINE bool wave_pass(X &x, SH &sh, const TableJob &J, const Scratch &W, Geo &G, const int box0[3][2],
                         const int flags0[6], int rad0, int radius, TeamCtl *tm = nullptr, int K = 1,
                         int member = 0, int *gen = nullptr) {
    const int T = x.nthr(), tid = x.tid();
    const bool lead = member == 0;
    const int nx = J.nx, ny = J.ny, nz = J.nz;
    const long long nxy = (long long)nx * ny, N = nxy * nz;
    const long long t0 = x.clock();
    WaveTabs WT;
    WT.mx = sh.arena;
    WT.my = WT.mx + nx;
    WT.mz = WT.my + ny;
    WT.ev = WT.mz + nz;
    WT.car = tm ? tm->car : WT.ev + (long long)kWaveEvInts * (nx + ny + nz);
    WT.pkb = tm ? tm->pk : WT.ev + 5;
    WT.pks = tm ? 1 : kWaveEvInts;
    // the rest of the arena groups a round's nodes by face type
    int *const lbuf = sh.arena + wave_arena_words(nx, ny, nz);
    const long long L = (long long)SH::kArena - wave_arena_words(nx, ny, nz);
    int *const wc = tm ? tm->wc : sh.wc;   // list sizes, requests waiting, error: shared by the team
    auto tsync = [&](bool light = false) -> bool {
        if (!tm) { x.sync(); return true; }
        return team_sync(x, tm, K, *gen, &sh.wc[kWcTeamOk], light);
    };
    // planes inside the starting box are interior, all others unswept so far
    for (int q = tid; q < nx + ny + nz; q += T) {
        int a, c;
        if (q < nx) { a = 0; c = q; } else if (q < nx + ny) { a = 1; c = q - nx; } else { a = 2; c = q - nx - ny; }
        WT.mx[q] = (c > box0[a][0] && c < box0[a][1]) ? -1 : kWaveNever;
    }
    for (int q = tid; q < 32; q += T) sh.wc[q] = 0;
    if (lead) {
        for (int q = tid; q < kWaveCarryCap * kWaveCarryInts; q += T) x.atomic_exch(&WT.car[q], 0);
        if (tm) {
            for (int q = tid; q < 16; q += T) x.atomic_exch(&tm->wc[q], 0);
            for (int q = tid; q < 8; q += T) x.atomic_exch(&tm->headw[q], 0);
            if (tid == 0) {
                x.atomic_exch(&tm->done, 0);
                x.atomic_exch(&tm->seeds, 0);
                x.atomic_exch(&tm->carry, 0);
            }
        }
    }
    x.sync();
    if (tid == 0) {
        // replay the controller (solve_table's first-pass loop) to list the events
        int x1 = box0[0][0], x2 = box0[0][1], y1 = box0[1][0], y2 = box0[1][1], z1 = box0[2][0], z2 = box0[2][1];
        int dx1 = flags0[0], dx2 = flags0[1], dy1 = flags0[2], dy2 = flags0[3], dz1 = flags0[4], dz2 = flags0[5];
        int ne = 0, total = 0, grow = rad0, rad = radius;
        auto rec = [&](int F, int plane) {
            int *E = WT.ev + kWaveEvInts * ne;
            E[0] = (int)(((unsigned)(x1 + 1)) | ((unsigned)(x2 + 1) << 16));
            E[1] = (int)(((unsigned)(y1 + 1)) | ((unsigned)(y2 + 1) << 16));
            E[2] = (int)(((unsigned)(z1 + 1)) | ((unsigned)(z2 + 1) << 16));
            E[3] = F;
            const int wa = (F == EAST || F == WEST) ? (y2 - y1 - 1) : (x2 - x1 - 1);
            const int wb = (F == TOP || F == BOTTOM) ? (y2 - y1 - 1) : (z2 - z1 - 1);
            E[4] = (wa > 0 && wb > 0) ? wa * wb : 0;
            E[5] = E[4];
            total += E[4];
            int *m = (F == TOP || F == BOTTOM) ? WT.mz : (F == EAST || F == WEST) ? WT.mx : WT.my;
            m[plane] = ne;
            ne++;
        };
        while (grow && (dx1 || dx2 || dy1 || dy2 || dz1 || dz2)) {
            for (int g = 1; g <= J.grow[0]; g++)
                if (dz1) { rec(TOP, z1); if (z1 == 0) dz1 = 0; z1--; }
            for (int g = 1; g <= J.grow[1]; g++)
                if (dz2) { rec(BOTTOM, z2); if (z2 == nz - 1) dz2 = 0; z2++; }
            for (int g = 1; g <= J.grow[2]; g++)
                if (dx2) { rec(EAST, x2); if (x2 == nx - 1) dx2 = 0; x2++; }
            for (int g = 1; g <= J.grow[3]; g++)
                if (dx1) { rec(WEST, x1); if (x1 == 0) dx1 = 0; x1--; }
            for (int g = 1; g <= J.grow[4]; g++)
                if (dy1) { rec(NORTH, y1); if (y1 == 0) dy1 = 0; y1--; }
            for (int g = 1; g <= J.grow[5]; g++)
                if (dy2) { rec(SOUTH, y2); if (y2 == ny - 1) dy2 = 0; y2++; }
            rad++;
            if (rad == J.maxrad) grow = 0;
        }
        sh.wc[kWcNev] = ne;
        sh.wc[kWcTotal] = total;
        if (lead) {
            // the slot's round numbers: every status word of earlier passes and
            // tables is below this pass's base (0: exhausted, start again at 1)
            unsigned b = *W.wround;
            if (b >= kWaveRoundLimit) b = 0;
            sh.wc[kWcDone] = (int)b;
        }
    }
    x.sync();
    unsigned base;
    if (lead) {
        base = (unsigned)sh.wc[kWcDone];
        if (base == 0) {
            for (long long q = tid; q < W.wn; q += T) W.wst[q] = 0u;
            base = 1;
        }
    } else {
        base = tm->base;
    }
    const int nev = sh.wc[kWcNev];
    const int stamp0 = lead ? sh.face_ctr + 1 : tm->stamp0;
    const float cin0 = sh.cin0, cin4 = sh.cin4;
    x.sync();
    if (tid == 0) sh.wc[kWcDone] = 0;
    if (lead && tm) {
        // open the pass for the other members
        for (int e = tid; e < nev; e += T) x.atomic_exch(&tm->pk[e], WT.fsize(e));
        if (tid == 0) {
            tm->W = W;
            for (int a = 0; a < 3; a++) { tm->box0[a][0] = box0[a][0]; tm->box0[a][1] = box0[a][1]; }
            for (int f = 0; f < 6; f++) tm->flags0[f] = flags0[f];
            tm->rad0 = rad0;
            tm->radius = radius;
            tm->stamp0 = stamp0;
            tm->base = base;
        }
        team_publish(x);
        if (tid == 0) x.store_release(&tm->seq, x.load_relaxed(&tm->seq) + 1);
    }
    x.sync();
    // counters: earlier-face predecessors of every node of the pass; nodes with
    // none and no in-face predecessor are ready for the first round (list wtmp)
    const WaveQ Q0{W.wtmp, &wc[kWcNCur], W.wcap, tm != nullptr};
    const long long first = (long long)member * T + tid, stride = (long long)K * T;
    for (long long v = first; v < N; v += stride) {
        int i, j, k;
        node_ijk(J, v, i, j, k);
        int ex[3], ey[3], ez[3];
        cube_events(J, WT, i, j, k, ex, ey, ez);
        const int e = imax3(ex[1], ey[1], ez[1]);
        if (e < 0 || e >= kWaveNever) continue;
        const int p1 = earlier_preds(ex, ey, ez, e);
        if (p1 > 0) {
            W.wcnt[v] = p1;
        } else {
            const int p2 = inface_preds(J, WT, i, j, k, e, ex, ey, ez);
            W.wcnt[v] = (p2 << 16) | (1 << 24);
            if (p2 == 0) {
                wave_queue(x, Q0, v, WT.type(e));
                x.atomic_add(&sh.st[kStatWaveSeeds], 1);
            }
        }
    }
    // the starting box's nodes are final keys of the faces beside it
    {
        const int lo[3] = {box0[0][0] + 1 > 0 ? box0[0][0] + 1 : 0, box0[1][0] + 1 > 0 ? box0[1][0] + 1 : 0,
                           box0[2][0] + 1 > 0 ? box0[2][0] + 1 : 0};
        const int hi[3] = {box0[0][1] - 1 < nx - 1 ? box0[0][1] - 1 : nx - 1,
                           box0[1][1] - 1 < ny - 1 ? box0[1][1] - 1 : ny - 1,
                           box0[2][1] - 1 < nz - 1 ? box0[2][1] - 1 : nz - 1};
        const long long wx = hi[0] - lo[0] + 1, wy = hi[1] - lo[1] + 1, wz = hi[2] - lo[2] + 1;
        const long long nin = (wx > 0 && wy > 0 && wz > 0) ? wx * wy * wz : 0;
        for (long long q = first; q < nin; q += stride) {
            const int i = lo[0] + (int)(q % wx), j = lo[1] + (int)((q / wx) % wy), k = lo[2] + (int)(q / (wx * wy));
            for (int d = 0; d < 6; d++) {
                const int ax = d >> 1, sg = (d & 1) ? 1 : -1;
                int ui = i, uj = j, uk = k;
                if (ax == 0) ui += sg; else if (ax == 1) uj += sg; else uk += sg;
                if (ui < 0 || uj < 0 || uk < 0 || ui >= nx || uj >= ny || uk >= nz) continue;
                const int eu = imax3(WT.mx[ui], WT.my[uj], WT.mz[uk]);
                if (eu < 0 || eu >= kWaveNever) continue;
                const int Fu = WT.type(eu);
                if (face_nax(Fu) == ax && face_ndir(Fu) == -sg) x.atomic_add(WT.pk(eu), -1);
            }
        }
    }
    // Rounds.  Round R evaluates list wtmp (R - base even) or wready (odd),
    // whose size is wc[kWcNCur + parity], and fills the other list and the
    // other size word.  Two barriers per round: after the evaluations, and
    // after the closing phase, during which the leader also resolves the
    // carried-register requests (no node value, status or key changes then)
    // and clears the size words the next round fills.
    unsigned R = base;
    bool ok = true, alive = tsync();   // alive: no member gave the pass up
    int idle = 0, why = 0;
    while (alive) {
        const long long tr0 = tid == 0 ? x.clock() : 0;
        const unsigned par = (R - base) & 1u;
        // read after the barrier
        const int ncur = x.load_relaxed(&wc[kWcNCur + par]), waiting = x.load_relaxed(&wc[kWcWaiting]);
        const int err = x.load_relaxed(&wc[kWcErr]);
        if (err || ncur > W.wcap) { ok = false; why = err ? 2 : 1; break; }
        if (ncur == 0 && waiting == 0) break;
        if (ncur == 0) {
            // only carried-register requests are open: one must resolve now
            if (++idle > 2) { ok = false; why = 3; break; }
        } else {
            idle = 0;
        }
        if (R - base > (unsigned)(2 * N + 64) || R >= (1u << 30) - 2) { ok = false; why = 4; break; }
        const int *cur = par ? W.wready : W.wtmp;
        const WaveQ Q{par ? W.wtmp : W.wready, &wc[kWcNCur + (1 - par)], W.wcap, tm != nullptr};
        const WaveQ CLQ{W.wclose, &wc[kWcNClose + par], W.wcap, true};
        const WaveQ *CL = tm ? &CLQ : nullptr;
        long long tr1 = 0;
        // this member's share of the list, grouped by face type in LDS so that
        // waves of one type evaluate together (the six evaluation routines are
        // large; waves running different ones at once miss in the instruction
        // cache); each evaluation queues the neighbours it leaves ready
        const long long c0 = (long long)ncur * member / K, c1 = (long long)ncur * (member + 1) / K;
        for (long long s0 = c0; s0 < c1; s0 += L) {
            const int m = (int)(c1 - s0 < L ? c1 - s0 : L);
            for (int f = tid; f < 6; f += T) sh.wc[kWcTypeCnt + f] = 0;
            x.sync();
            for (int it = tid; it < m; it += T)
                x.atomic_add(&sh.wc[kWcTypeCnt + (int)((unsigned)wave_entry(x, cur, s0 + it, tm) >> 29)], 1);
            x.sync();
            if (tid == 0) {
                int o = 0;
                for (int f = 0; f < 6; f++) {
                    sh.wc[kWcTypeOff + f] = o;
                    o += sh.wc[kWcTypeCnt + f];
                }
            }
            x.sync();
            for (int it = tid; it < m; it += T) {
                const int w = wave_entry(x, cur, s0 + it, tm);
                lbuf[x.atomic_add(&sh.wc[kWcTypeOff + (int)((unsigned)w >> 29)], 1)] = w;
            }
            x.sync();
            if (tid == 0 && tr1 == 0) tr1 = x.clock();
            // each type's offset now ends its range
            static_for<0, 6>([&](auto FT) {
                constexpr int Fc = decltype(FT)::value;
                const int lo = Fc == 0 ? 0 : sh.wc[kWcTypeOff + Fc - 1], hi = sh.wc[kWcTypeOff + Fc];
                if (lo == hi) return;   // uniform: offsets read after the barrier
                for (int it = lo + tid; it < hi; it += T)
                    wave_eval<Fc>(x, sh, J, W, G, WT, Q, CL, lbuf[it] & 0x1fffffff, base, R, stamp0);
                x.sync();
            });
        }
        if (lead && tid == 0) {
            sh.st[kStatRoundsWave] += 1;
            sh.st[kStatEvals] += ncur;
        }
        if (!(alive = tsync())) break;
        const long long tr2 = tid == 0 ? x.clock() : 0;
        // closing phase: every key written this round is visible now
        const int ncl = x.load_relaxed(&wc[kWcNClose + par]);
        if (ncl > W.wcap) { ok = false; why = 5; break; }   // uniform: read after the barrier
        for (long long it = first; it < ncl; it += stride) wave_close(x, J, W, WT, Q, wave_entry(x, W.wclose, it, true));
        if (lead) {
            if (tid == 0) {
                // every member has read this round's size: the round after
                // next fills that word; the next round fills the other
                // closing word
                x.atomic_exch(&wc[kWcNCur + par], 0);
                x.atomic_exch(&wc[kWcNClose + (1 - par)], 0);
            }
            // carried-register requests: find each one's predecessor (a
            // failure is seen by every member when the next round starts)
            bool fail = false;
            for (int c = 0; c < kWaveCarryCap && !fail; c++) {
                int *q = WT.car + kWaveCarryInts * c;
                if (x.load_relaxed(&q[5]) != 1) continue;   // uniform: no member changes requests now
                for (int guard = 0; guard < 4 * nev + 8; guard++) {
                    const int cand = x.load_relaxed(&q[4]), qv = x.load_relaxed(&q[0]), qe = x.load_relaxed(&q[1]);
                    if (cand >= 0) {
                        const unsigned s = W.wst[cand];
                        if (!wv_final_in(s, base)) break;            // wait for it
                        if (s & 2u) {                                // it set them: its t[0]/t[4]
                            x.sync();
                            if (tid == 0) {
                                x.store_relaxed((int *)&W.wt0[qv], __builtin_bit_cast(int, W.wt0[cand]));
                                x.store_relaxed((int *)&W.wt4[qv], __builtin_bit_cast(int, W.wt4[cand]));
                                x.atomic_exch(&q[5], 2);
                                wave_queue(x, Q, qv, WT.type(qe));
                            }
                            x.sync();
                            break;
                        }
                        x.sync();
                        if (tid == 0) x.atomic_exch(&q[4], -1);   // it passed its own carry on: search below it
                        x.sync();
                        continue;
                    }
                    const int e = x.load_relaxed(&q[6]);
                    if (e < 0) {   // before the first face: the pass's carry-in
                        x.sync();
                        if (tid == 0) {
                            x.store_relaxed((int *)&W.wt0[qv], __builtin_bit_cast(int, cin0));
                            x.store_relaxed((int *)&W.wt4[qv], __builtin_bit_cast(int, cin4));
                            x.atomic_exch(&q[5], 2);
                            wave_queue(x, Q, qv, WT.type(qe));
                        }
                        x.sync();
                        break;
                    }
                    const unsigned long long bound = ((unsigned long long)(unsigned)x.load_relaxed(&q[3]) << 32) |
                                                     (unsigned)x.load_relaxed(&q[2]);
                    if (WT.fsize(e) > 0 && x.load_relaxed(WT.pk(e)) != 0) break;   // keys of that face not all final yet
                    unsigned long long best = ~0ull;
                    if (WT.fsize(e) > 0) {
                        best = wave_scan(x, sh, J, W, WT, e, bound, false, base);
                        if (tid == 0) sh.st[kStatWaveScans] += 1;
                        if (sh.wc[kWcScanBad]) {   // a key not final although its face count says so
                            fail = true;
                            break;
                        }
                    }
                    x.sync();
                    if (tid == 0) {
                        if (best == ~0ull) {   // continue at the last node of the previous face
                            x.atomic_exch(&q[6], e - 1);
                            x.atomic_exch(&q[2], -1);
                            x.atomic_exch(&q[3], -1);
                        } else {
                            FaceGeom fg;
                            wave_face_geom(WT, e, fg);
                            int ci, cj, ck;
                            face_coords(fg, (int)(best & 0xffffffffu), ci, cj, ck);
                            x.atomic_exch(&q[4], (int)((long long)ck * nxy + (long long)cj * nx + ci));
                            x.atomic_exch(&q[2], (int)(unsigned)(best & 0xffffffffu));
                            x.atomic_exch(&q[3], (int)(unsigned)(best >> 32));
                        }
                    }
                    x.sync();
                }
            }
            x.sync();
            if (tid == 0) {
                int nw = 0;
                for (int c = 0; c < kWaveCarryCap; c++) nw += x.load_relaxed(&WT.car[kWaveCarryInts * c + 5]) != 0;
                x.atomic_exch(&wc[kWcWaiting], nw);
                if (fail) x.atomic_exch(&wc[kWcErr], 1);
            }
        }
        // what this phase wrote (list entries and sizes, counters, requests,
        // resolved carry-ins) went through atomics: an ordering barrier
        if (!(alive = tsync(true))) break;
        if (tid == 0 && lead) {
            const long long tr3 = x.clock();
            sh.st[kStatTimeWaveGroup] += (int)((tr1 ? tr1 : tr2) - tr0);
            sh.st[kStatTimeWaveEval] += (int)(tr2 - (tr1 ? tr1 : tr2));
            sh.st[kStatTimeWaveCarry] += (int)(tr3 - tr2);
        }
        R++;
    }
    // every member leaves the rounds at the same barrier; the members' counts
    // go to the leader
    if (tm && alive) {
        if (tid == 0) {
            x.atomic_add(&tm->done, sh.wc[kWcDone]);
            if (!lead) {
                for (int w = 0; w < 7; w++)
                    if (sh.headw[w]) x.atomic_add(&tm->headw[w], sh.headw[w]);
                x.atomic_add(&tm->seeds, sh.st[kStatWaveSeeds]);
                x.atomic_add(&tm->carry, sh.st[kStatWaveCarry]);
                sh.st[kStatWaveSeeds] = 0;
                sh.st[kStatWaveCarry] = 0;
            }
        }
        alive = tsync();
    }
    if (!alive) { ok = false; why = 7; }
    if (!lead) return ok;
    if (tm && !alive) {
        // a member gave up: wait until the others have left the slot's state
        if (tid == 0) {
            while (x.atomic_add(&tm->inpass, 0) != 0) x.pause();
            x.fence_acquire();
        }
        x.sync();
    }
    // every node of the pass must have been evaluated exactly once
    if (ok && (tm ? x.load_relaxed(&tm->done) : sh.wc[kWcDone]) != sh.wc[kWcTotal]) { ok = false; why = 6; }
    if (tid == 0) {
        if (tm && alive) {
            for (int w = 0; w < 7; w++) sh.headw[w] += x.load_relaxed(&tm->headw[w]);
            sh.st[kStatWaveSeeds] += x.load_relaxed(&tm->seeds);
            sh.st[kStatWaveCarry] += x.load_relaxed(&tm->carry);
        }
        if (!ok) {
            sh.st[kStatError] |= 128;
            sh.st[kStatWaveFail] = why;
        }
    }
    x.sync();
    if (ok) {
        // t[0]/t[4] leaving the pass: those of the last node in sequential order
        // that set them (unchanged if no node did)
        for (int e = nev - 1; e >= 0; e--) {
            if (WT.fsize(e) == 0) continue;
            const unsigned long long best = wave_scan(x, sh, J, W, WT, e, ~0ull, true, base);
            if (best == ~0ull) continue;
            if (tid == 0) {
                FaceGeom fg;
                wave_face_geom(WT, e, fg);
                int ci, cj, ck;
                face_coords(fg, (int)(best & 0xffffffffu), ci, cj, ck);
                const long long u = (long long)ck * nxy + (long long)cj * nx + ci;
                sh.cin0 = W.wt0[u];
                sh.cin4 = W.wt4[u];
            }
            break;
        }
    }
    if (tid == 0) {
        *W.wround = R + 1;
        sh.face_ctr += nev;
        sh.st[kStatWavePasses] += 1;
        sh.st[kStatTimeWave] += (int)(x.clock() - t0);
    }
    x.sync();
    return ok;
}

// The grid's row and layer tables into LDS (filled in J.geo_fill first when
// the host sent no tables) and the stencils' geometry constants.
template <class X, class SH>
FD_INLINE void load_geo(X &x, SH &sh, const TableJob &J, Geo &G) {
    const double *gbuf = J.geo_fill ? J.geo_fill : J.geo;
    const long long nrl = geo_rows_doubles(J.ny) + geo_layers_doubles(J.nz);
    for (long long q = x.tid(); q < nrl; q += x.nthr()) sh.geol[q] = gbuf[q];
    G.h = J.h; G.dq = J.dq; G.df = J.df; G.y0 = J.y0; G.z0r = J.z0r; G.sinq = J.sinq;
    G.rh = 1.0 / J.h;
    G.ihsq = 1. / (J.h * J.h);
    G.idqsq = 1. / (J.dq * J.dq);
    G.rowt = (GeoPtr)sh.geol;
    G.layt = G.rowt + geo_rows_doubles(J.ny);
    G.nx = J.nx; G.ny = J.ny; G.nz = J.nz;
    x.sync();
}

// Member `member` (> 0) of the team solving table J: joins every wavefront
// pass the leader opens, with the leader's slot, until the leader releases
// the team.  Its own slot is not used.
template <class X, class SH>
FD_HD void team_helper(X &x, SH &sh, const TableJob &J, TeamCtl *tm, int K, int member) {
    const int tid = x.tid();
    int seen = 0, gen = 0;
    bool geo = false;
    Geo G;
    for (;;) {
        if (tid == 0) {
            // wait for the next pass (or the end of the table)
            int go = 0;
            const long long w0 = x.clock(), lim = tm->wait_ticks > 0 ? tm->wait_ticks : kTeamWaitTicks;
            for (;;) {
                if (x.load_relaxed(&tm->abort) || x.load_relaxed(&tm->release)) break;
                if (x.load_relaxed(&tm->seq) > seen) {
                    // inside the pass unless a member gave up meanwhile (the
                    // leader waits for inpass 0 before it reuses its slot)
                    x.atomic_add(&tm->inpass, 1);
                    if (x.atomic_add(&tm->abort, 0)) x.atomic_add(&tm->inpass, -1);
                    else go = 1;
                    break;
                }
                if (x.clock() - w0 > lim) { x.atomic_exch(&tm->abort, 1); break; }
                x.pause();
            }
            x.fence_acquire();
            sh.wc[kWcTeamOk] = go;
        }
        x.sync();
        const int go = sh.wc[kWcTeamOk];
        x.sync();
        if (!go) break;
        seen++;
        if (!geo) {
            // the leader's checks passed: its geometry tables fit the arena
            load_geo(x, sh, J, G);
            geo = true;
        }
        const Scratch W = tm->W;
        int box0[3][2], flags0[6];
        for (int a = 0; a < 3; a++) { box0[a][0] = tm->box0[a][0]; box0[a][1] = tm->box0[a][1]; }
        for (int f = 0; f < 6; f++) flags0[f] = tm->flags0[f];
        const int rad0 = tm->rad0, radius = tm->radius;
        if (tid == 0) {
            for (int w = 0; w < 7; w++) sh.headw[w] = 0;
            for (int q = 0; q < kStatCount; q++) sh.st[q] = 0;
        }
        x.sync();
        wave_pass(x, sh, J, W, G, box0, flags0, rad0, radius, tm, K, member, &gen);
        team_publish(x);
        if (tid == 0) x.atomic_add(&tm->inpass, -1);
        x.sync();
    }
    x.sync();
}

}  // namespace fdsweep
