// Host-thread executor for the sweep solver (fdtomoc_amd/csrc/sweep_solver.h).
//
// Test infrastructure: runs the exact bulk-synchronous code that the gfx950
// kernel runs, with std::thread workers standing in for the lanes of one
// workgroup, so the schedule (LEVEL / CHANGE, carried-register resolution,
// radix sort) can be checked against the oracle on a machine without a GPU.
// It is not part of the product and is never loaded by it.
#include <algorithm>
#include <barrier>
#include <chrono>
#include <memory>
#include <cstring>
#include <thread>
#include <vector>

#include "../../fdtomoc_amd/csrc/sweep_solver.h"
#include "../../fdtomoc_amd/csrc/table_prep.h"

namespace {
struct HostExec {
    int t, n;
    std::barrier<> *bar;
    int tid() const { return t; }
    int nthr() const { return n; }
    void sync() { bar->arrive_and_wait(); }
    int atomic_add(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
    int atomic_exch(int *p, int v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
    unsigned atomic_sub_u(unsigned *p, unsigned v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
    unsigned atomic_or(unsigned *p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
    unsigned atomic_and(unsigned *p, unsigned v) { return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST); }
    unsigned atomic_cas(unsigned *p, unsigned cmp, unsigned v) {
        __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
        return cmp;
    }
    unsigned long long atomic_max_u64(unsigned long long *p, unsigned long long v) {
        unsigned long long o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
        while (o < v && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
        }
        return o;
    }
    long long clock() const {
        return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
                   .count() / 10;
    }
    void fence_release() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
    void fence_acquire() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
    int load_relaxed(const int *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
    void store_relaxed(int *p, int v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }
    void wait_stores() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
    void store_release(int *p, int v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
    void pause() { std::this_thread::yield(); }
};
}  // namespace

// Same layout as oracle_par (oracle/sphfd_oracle.h) so one ctypes struct serves both.
struct EmuPar {
    int nx, ny, nz;
    double h, dq, df, x0, y0, z0;
    int iminus, iplus, jminus, jplus, kminus, kplus;
    int floatsrc, srctype, srcwall;
    double fxs, fys, fzs;
    int xs, ys, zs;
    int ncube, reverse, headpref, parse, fill;
    float maxoff;
    int invh, invq, invf, invx0, invy0, invz0;
};

// Workgroup state whose schedules keep no per-node state in the arena: every
// LEVEL face takes its pending counters and successor masks from global
// scratch (Scratch::pend, Scratch::flag) and every CHANGE face its queue marks
// (Scratch::stamp), the paths the device takes for faces too large for LDS.
using SharedGlobal = fdsweep::SharedT<fdsweep::kArenaWords, fdsweep::kGeoLdsDoubles, 0>;

// One table on `team` workgroups of nthreads threads each: workgroup 0 solves
// it (the leader), the others join its wavefront passes as the device's team
// members do (team 1: a lone workgroup).
template <class SH>
static void run_threads(int nthreads, int team, const fdsweep::TableJob &J, const fdsweep::Scratch &W,
                        bool missing = false) {
    std::vector<std::unique_ptr<SH>> shs;
    std::vector<std::unique_ptr<std::barrier<>>> bars;
    for (int g = 0; g < team; g++) {
        shs.emplace_back(new SH());
        std::memset(shs.back().get(), 0, sizeof(SH));
        bars.emplace_back(new std::barrier<>(nthreads));
    }
    std::unique_ptr<fdsweep::TeamCtl> tm(new fdsweep::TeamCtl());
    std::memset(tm.get(), 0, sizeof(fdsweep::TeamCtl));
    // a missing member: the others give the pass up after 20 ms
    tm->wait_ticks = missing ? 2000000 : 0;
    std::vector<std::thread> th;
    for (int g = 0; g < team - (missing ? 1 : 0); g++)
        for (int t = 0; t < nthreads; t++)
            th.emplace_back([&, g, t]() {
                HostExec x{t, nthreads, bars[(size_t)g].get()};
                SH &sh = *shs[(size_t)g];
                if (team == 1) {
                    fdsweep::solve_table(x, sh, J, W);
                } else if (g == 0) {
                    fdsweep::solve_table(x, sh, J, W, tm.get(), team);
                    fdsweep::team_release(x, tm.get());
                } else {
                    fdsweep::team_helper(x, sh, J, tm.get(), team, g);
                }
            });
    for (auto &t : th) t.join();
}

// mode: bit0 per-node schedule state always in global scratch (else in the
// arena where a face fits, as the device builds); bit1 the first pass as a
// wavefront, bit2 reverse passes as wavefronts; bits 3-7 the number of
// workgroups sharing the wavefront passes (0 or 1: one); bit 8 the team's last
// member never starts (its failure path).
extern "C" int emu_solve(const EmuPar *pc, const float *slow, float *time, const float *extra, int nthreads,
                         int *stats, int change_budget, int mode) {
    const int global_state = mode & 1;
    fdprep::ParValues p;
    p.nx = pc->nx; p.ny = pc->ny; p.nz = pc->nz;
    p.h = pc->h; p.dq = pc->dq; p.df = pc->df; p.x0 = pc->x0; p.y0 = pc->y0; p.z0 = pc->z0;
    p.iminus = pc->iminus; p.iplus = pc->iplus; p.jminus = pc->jminus; p.jplus = pc->jplus;
    p.kminus = pc->kminus; p.kplus = pc->kplus;
    p.floatsrc = pc->floatsrc; p.srctype = pc->srctype; p.srcwall = pc->srcwall;
    p.fxs = pc->fxs; p.fys = pc->fys; p.fzs = pc->fzs; p.xs = pc->xs; p.ys = pc->ys; p.zs = pc->zs;
    p.ncube = pc->ncube; p.reverse = pc->reverse; p.headpref = pc->headpref; p.parse = pc->parse; p.fill = pc->fill;
    p.maxoff = pc->maxoff;
    p.invh = pc->invh; p.invq = pc->invq; p.invf = pc->invf; p.invx0 = pc->invx0; p.invy0 = pc->invy0; p.invz0 = pc->invz0;
    fdprep::Prepared pr;
    int rc = fdprep::prepare_geometry(p, pr);
    if (rc) return rc;
    const bool preset = p.srctype != 1;
    rc = fdprep::prepare_source(p, pr, preset ? extra : nullptr, (p.fill == 1) ? extra : nullptr, preset ? time : nullptr);
    if (rc) return rc;
    const long long maxF = fdprep::max_face(p.nx, p.ny, p.nz);
    std::vector<float> O(maxF), tout0(maxF), tout4(maxF);
    std::vector<unsigned> key(maxF), key2(maxF);
    std::vector<int> val(maxF), val2(maxF), rank(maxF), ord(maxF), pend(maxF), stat(maxF), l0(maxF),
        l1(maxF), l2(maxF), stamp(maxF);
    std::vector<unsigned char> flag(maxF);
    fdsweep::Scratch W;
    W.O = O.data(); W.key = key.data(); W.key2 = key2.data(); W.val = val.data(); W.val2 = val2.data();
    W.rank = rank.data(); W.ord = ord.data(); W.pend = pend.data(); W.stat = stat.data();
    W.tout0 = tout0.data(); W.tout4 = tout4.data();
    W.list0 = l0.data(); W.list1 = l1.data(); W.list2 = l2.data(); W.stamp = stamp.data(); W.flag = flag.data();
    W.mstamp = nullptr; W.lastf = nullptr;
    const long long nn = (long long)p.nx * p.ny * p.nz;
    std::vector<unsigned> wst((size_t)nn, 0u);
    std::vector<float> wt0((size_t)nn), wt4((size_t)nn);
    std::vector<int> wl((size_t)nn), wtmp((size_t)nn), wrd((size_t)nn), wcl((size_t)nn);
    unsigned wround = 1;
    W.wst = wst.data(); W.wt0 = wt0.data(); W.wt4 = wt4.data();
    W.wcnt = wl.data(); W.wtmp = wtmp.data(); W.wready = wrd.data(); W.wclose = wcl.data();
    W.wcap = nn; W.wn = nn; W.wround = &wround;
    fdsweep::TableJob J;
    std::memset(&J, 0, sizeof J);
    J.nx = p.nx; J.ny = p.ny; J.nz = p.nz;
    J.h = pr.h; J.dq = pr.dq; J.df = pr.df; J.y0 = pr.y0; J.z0r = pr.z0r; J.sinq = pr.sinq.data();
    J.slow = slow; J.s_sy = p.nx; J.s_sz = (long long)p.nx * p.ny; J.time = time;
    J.init_mode = pr.init_mode; J.src_sidx = pr.src_index; J.ninit = (int)pr.init_idx.size();
    J.init_idx = pr.init_idx.data(); J.init_dist = pr.init_dist.data(); J.init_val = pr.init_val.data();
    std::memcpy(J.box, pr.box, sizeof J.box);
    std::memcpy(J.dflag, pr.dflag, sizeof J.dflag);
    std::memcpy(J.grow, pr.grow, sizeof J.grow);
    J.reverse = pr.reverse; J.headpref = pr.headpref; J.maxrad = pr.maxrad; J.radius0 = pr.radius0;
    J.srcwall = pr.srcwall; J.stats = stats;
    const int maxdim = std::max(p.nx, std::max(p.ny, p.nz));
    std::vector<int> mod_stamp((size_t)p.nx * p.ny * p.nz), last_full(6 * (size_t)maxdim);
    J.mod_stamp = mod_stamp.data();
    J.last_full = last_full.data();
    J.maxdim = maxdim;
    float smin = 1e30f, smax = 0.f;
    for (long long q = 0; q < (long long)p.nx * p.ny * p.nz; q++) {
        smin = slow[q] < smin ? slow[q] : smin;
        smax = slow[q] > smax ? slow[q] : smax;
    }
    J.skip_ok = fdprep::skip_safe(p, pr, 1.0 / smax, 1.0 / smin) ? 1 : 0;
    J.change_budget = change_budget;
    J.wave = (mode >> 1) & 3;
    std::vector<double> geo((size_t)fdsweep::geo_doubles(p.ny, p.nz));
    J.geo_fill = geo.data();
    const int team = std::max(1, std::min(fdsweep::kTeamMax, (mode >> 3) & 31));
    const bool missing = team > 1 && ((mode >> 8) & 1);
    if (global_state) run_threads<SharedGlobal>(nthreads, team, J, W, missing);
    else run_threads<fdsweep::Shared>(nthreads, team, J, W, missing);
    return 0;
}
