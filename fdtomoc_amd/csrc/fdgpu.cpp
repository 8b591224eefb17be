// fdgpu.cpp -- host side of the batched/served sphfd GPU API (include/fdtomoc_amd.h).
//
// Owns device memory (models, resident tables, per-workgroup scratch),
// prepares each table on the host (table_prep.cpp: the reference's geometry and
// source set-up) and launches the persistent sweep kernel.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "../../include/fdtomoc_amd.h"
#include "fd_kernels.h"
#include "fdgpu_internal.h"
#include "sphfd_io.h"
#include "table_prep.h"

namespace fdgpu_detail {

int hip_fail(hipError_t e) { return e == hipErrorOutOfMemory ? FDGPU_ENOMEM : FDGPU_EHIP; }

#define FD_HIP(call)                                  \
    do {                                              \
        hipError_t _e = (call);                       \
        if (_e != hipSuccess) return hip_fail(_e);    \
    } while (0)

int dev_alloc(fdgpu_ctx *c, void **p, size_t bytes) {
    if (c->used + bytes > c->budget) return FDGPU_ENOMEM;
    FD_HIP(hipMalloc(p, bytes ? bytes : 1));
    c->used += bytes;
    return FDGPU_OK;
}

void dev_free(fdgpu_ctx *c, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    c->used -= bytes;
}

// Releases the launch slots (regrown by the next launch).
void release_scratch(fdgpu_ctx *c) {
    if (!c->d_scratch && !c->d_slots) return;
    (void)hipStreamSynchronize(c->stream);
    clear_graphs(c);
    dev_free(c, c->d_scratch, c->scratch_bytes);
    dev_free(c, c->d_slots, sizeof(fdsweep::Scratch) * (size_t)c->slots_cap);
    c->d_scratch = nullptr;
    c->d_slots = nullptr;
    c->scratch_bytes = 0;
    c->slots_cap = 0;
    c->slot_face_cap = 0;
    c->slot_node_cap = 0;
    c->slot_dim_cap = 0;
    c->slot_wave = false;
}

// The ray tracer's cached device arrays (regrown by its next call).
void release_ray_columns(fdgpu_ctx *c) {
    if (!c->ray_cols && !c->ray_blk && !c->ray_out) return;
    (void)hipStreamSynchronize(c->stream);
    dev_free(c, c->ray_cols, c->ray_cols_bytes);
    dev_free(c, c->ray_blk, c->ray_blk_bytes);
    dev_free(c, c->ray_out, c->ray_out_bytes);
    c->ray_cols = nullptr;
    c->ray_cols_bytes = 0;
    c->ray_blk = nullptr;
    c->ray_blk_bytes = 0;
    c->ray_out = nullptr;
    c->ray_out_bytes = 0;
}

// Memory that outlives a launch (tables, models): when the budget is short,
// the launch slots and the ray tracer's columns are released first; the
// next launch or trace regrows them within what is left.
int dev_alloc_resident(fdgpu_ctx *c, void **p, size_t bytes) {
    int rc = dev_alloc(c, p, bytes);
    if (rc == FDGPU_ENOMEM && (c->d_scratch || c->ray_cols || c->ray_blk || c->ray_out)) {
        release_scratch(c);
        release_ray_columns(c);
        rc = dev_alloc(c, p, bytes);
    }
    return rc;
}

// Capacity of each per-round list of the wavefront passes: a round queues a
// surface of the table, far below this; a pass that would overflow it
// reports it and the table is solved face by face instead.
long long wave_list_cap(long long N) { return std::min(N, (long long)1 << 22); }

size_t slot_bytes(long long F, long long N, int D, bool wave) {
    const size_t f = align256(sizeof(float) * (size_t)F);
    size_t b = 15 * f + align256((size_t)F);
    if (N > 0) {
        b += align256(sizeof(int) * (size_t)N) + align256(sizeof(int) * 6 * (size_t)D);
        // wavefront state: status, t[0], t[4], counter per node, three lists, the round word
        if (wave)
            b += 4 * align256(sizeof(int) * (size_t)N) + 3 * align256(sizeof(int) * (size_t)wave_list_cap(N)) + 256;
    }
    return b;
}

int fill_slots(char *base, int ns, long long F, fdsweep::Scratch *d_slots, long long N, int D, bool wave) {
    const size_t per = slot_bytes(F, N, D, wave);
    const size_t f = align256(sizeof(float) * (size_t)F);
    std::vector<fdsweep::Scratch> h((size_t)ns);
    for (int s = 0; s < ns; s++) {
        char *b = base + per * (size_t)s;
        fdsweep::Scratch &W = h[(size_t)s];
        size_t o = 0;
        auto take = [&](size_t n) { char *p = b + o; o += n; return p; };
        W.O = (float *)take(f);
        W.key = (unsigned *)take(f);
        W.key2 = (unsigned *)take(f);
        W.val = (int *)take(f);
        W.val2 = (int *)take(f);
        W.rank = (int *)take(f);
        W.ord = (int *)take(f);
        W.pend = (int *)take(f);
        W.stat = (int *)take(f);
        W.tout0 = (float *)take(f);
        W.tout4 = (float *)take(f);
        W.list0 = (int *)take(f);
        W.list1 = (int *)take(f);
        W.list2 = (int *)take(f);
        W.stamp = (int *)take(f);
        W.flag = (unsigned char *)take(align256((size_t)F));
        W.mstamp = nullptr;
        W.lastf = nullptr;
        W.wst = nullptr; W.wt0 = nullptr; W.wt4 = nullptr;
        W.wcnt = nullptr; W.wtmp = nullptr; W.wready = nullptr; W.wclose = nullptr;
        W.wcap = 0; W.wn = 0; W.wround = nullptr;
        if (N > 0) {
            W.mstamp = (int *)take(align256(sizeof(int) * (size_t)N));
            W.lastf = (int *)take(align256(sizeof(int) * 6 * (size_t)D));
        }
        if (N > 0 && wave) {
            const size_t nb = align256(sizeof(int) * (size_t)N), lb = align256(sizeof(int) * (size_t)wave_list_cap(N));
            W.wst = (unsigned *)take(nb);
            W.wt0 = (float *)take(nb);
            W.wt4 = (float *)take(nb);
            W.wcnt = (int *)take(nb);
            W.wtmp = (int *)take(lb);
            W.wready = (int *)take(lb);
            W.wclose = (int *)take(lb);
            W.wcap = wave_list_cap(N);
            W.wn = N;
            W.wround = (unsigned *)take(256);
            // round word 0: the slot's first wavefront pass clears its status words
            FD_HIP(hipMemset(W.wround, 0, sizeof(unsigned)));
        }
    }
    FD_HIP(hipMemcpy(d_slots, h.data(), sizeof(fdsweep::Scratch) * (size_t)ns, hipMemcpyHostToDevice));
    return FDGPU_OK;
}

// Slots for a launch: face scratch for faces of up to maxF nodes and the
// change stamps of tables of up to maxN nodes (the stamps live only for the
// duration of one table's solve, so they belong to the slot, not the table);
// grown on demand, never shrunk.
int ensure_scratch(fdgpu_ctx *c, int nslots, long long maxF, long long maxN, int maxD, bool wave) {
    const bool fits = nslots <= c->slots_cap && maxF <= c->slot_face_cap && maxN <= c->slot_node_cap &&
                      maxD <= c->slot_dim_cap;
    if (fits && (c->slot_wave || !wave)) return FDGPU_OK;
    const int ns = std::max(nslots, c->slots_cap);
    const long long F = std::max(maxF, c->slot_face_cap);
    const long long N = std::max(maxN, c->slot_node_cap);
    const int D = std::max(maxD, c->slot_dim_cap);
    // wavefront state only where it takes at most half of what the budget
    // has left beside everything else (the slots' own earlier allocation is
    // released first), so that it does not crowd out the tables a caller
    // keeps resident
    if (wave) {
        const size_t avail = c->budget - (c->used - c->scratch_bytes - sizeof(fdsweep::Scratch) * (size_t)c->slots_cap);
        wave = slot_bytes(F, N, D, true) * (size_t)ns + sizeof(fdsweep::Scratch) * (size_t)ns <= avail / 2;
        if (!wave && fits) return FDGPU_OK;
    }
    release_scratch(c);
    const size_t per = slot_bytes(F, N, D, wave);
    int rc = dev_alloc(c, (void **)&c->d_scratch, per * (size_t)ns);
    if (rc) return rc;
    c->scratch_bytes = per * (size_t)ns;
    rc = dev_alloc(c, (void **)&c->d_slots, sizeof(fdsweep::Scratch) * (size_t)ns);
    if (rc) return rc;
    rc = fill_slots(c->d_scratch, ns, F, c->d_slots, N, D, wave);
    if (rc) return rc;
    c->slot_wave = wave;
    c->slots_cap = ns;
    c->slot_face_cap = F;
    c->slot_node_cap = N;
    c->slot_dim_cap = D;
    return FDGPU_OK;
}

fdprep::ParValues to_values(const fdgpu_par &p) {
    fdprep::ParValues v;
    v.nx = p.nx; v.ny = p.ny; v.nz = p.nz;
    v.h = p.h; v.dq = p.dq; v.df = p.df; v.x0 = p.x0; v.y0 = p.y0; v.z0 = p.z0;
    v.iminus = p.iminus; v.iplus = p.iplus; v.jminus = p.jminus; v.jplus = p.jplus; v.kminus = p.kminus; v.kplus = p.kplus;
    v.floatsrc = p.floatsrc; v.srctype = p.srctype; v.srcwall = p.srcwall;
    v.fxs = p.fxs; v.fys = p.fys; v.fzs = p.fzs; v.xs = p.xs; v.ys = p.ys; v.zs = p.zs;
    v.ncube = p.ncube; v.reverse = p.reverse; v.headpref = p.headpref; v.parse = p.parse; v.fill = p.fill;
    v.maxoff = p.maxoff;
    v.invh = p.invh; v.invq = p.invq; v.invf = p.invf; v.invx0 = p.invx0; v.invy0 = p.invy0; v.invz0 = p.invz0;
    v.swab = p.swab;
    return v;
}

void clear_geos(fdgpu_ctx *c) {
    for (auto &kv : c->geos) dev_free(c, kv.second.d, kv.second.bytes);
    c->geos.clear();
}

namespace {

unsigned long long dbits(double v) {
    unsigned long long u;
    memcpy(&u, &v, 8);
    return u;
}

// The geometry table of a table's grid from the cache, computed on the
// device on first use (J.sinq must already be on the device).  When the
// budget is short, cached tables that `keep` does not name are dropped first.
int geometry_for(fdgpu_ctx *c, const fdsweep::TableJob &J, const std::vector<double> &sinq,
                 std::map<std::vector<unsigned long long>, int> &keep, const double **out) {
    std::vector<unsigned long long> key = {(unsigned long long)J.ny, (unsigned long long)J.nz, dbits(J.h), dbits(J.dq),
                                           dbits(J.df), dbits(J.z0r)};
    for (double v : sinq) key.push_back(dbits(v));
    auto it = c->geos.find(key);
    if (it == c->geos.end()) {
        GeoBuf g;
        g.bytes = sizeof(double) * (size_t)fdsweep::geo_doubles(J.ny, J.nz);
        int rc = dev_alloc(c, (void **)&g.d, g.bytes);
        if (rc == FDGPU_ENOMEM) {
            (void)hipStreamSynchronize(c->stream);
            for (auto g2 = c->geos.begin(); g2 != c->geos.end();) {
                if (keep.count(g2->first)) { ++g2; continue; }
                dev_free(c, g2->second.d, g2->second.bytes);
                g2 = c->geos.erase(g2);
            }
            rc = dev_alloc(c, (void **)&g.d, g.bytes);
        }
        if (rc) return rc;
        const fdsweep::GeoParams gp{J.ny, J.nz, J.h, J.dq, J.df, J.z0r, J.sinq};
        if (fd_launch_geo(&gp, g.d, c->stream) != hipSuccess) {
            dev_free(c, g.d, g.bytes);
            return FDGPU_EHIP;
        }
        it = c->geos.emplace(key, g).first;
    }
    keep[key] = 1;
    *out = it->second.d;
    return FDGPU_OK;
}

}  // namespace

void clear_graphs(fdgpu_ctx *c) {
    for (auto &kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
}

// Device array of table descriptors, kept between solves so that captured
// launches stay valid; grown on demand.
int ensure_jobs(fdgpu_ctx *c, int n) {
    if (n <= c->jobs_cap) return FDGPU_OK;
    (void)hipStreamSynchronize(c->stream);
    clear_graphs(c);
    dev_free(c, c->d_jobs, sizeof(fdsweep::TableJob) * (size_t)c->jobs_cap);
    c->d_jobs = nullptr;
    c->jobs_cap = 0;
    const int cap = std::max(n, 2 * c->jobs_cap);
    int rc = dev_alloc(c, (void **)&c->d_jobs, sizeof(fdsweep::TableJob) * (size_t)cap);
    if (rc) return rc;
    c->jobs_cap = cap;
    return FDGPU_OK;
}

// Tables below this many nodes are never solved by a team: their rounds hold
// too few nodes to share.
constexpr long long kTeamMinNodes = 1LL << 20;

// Team blocks and the teams' first tickets (cap + 1 ints) in one allocation.
size_t team_bytes(int cap) {
    return cap > 0 ? sizeof(fdsweep::TeamCtl) * (size_t)cap + sizeof(int) * (size_t)(cap + 1) : 0;
}

// Team blocks for nteam teams (the captured launches keep their address, so
// growing them drops the graphs).
int ensure_teams(fdgpu_ctx *c, int nteam) {
    if (nteam <= c->teams_cap) return FDGPU_OK;
    (void)hipStreamSynchronize(c->stream);
    clear_graphs(c);
    dev_free(c, c->d_teams, team_bytes(c->teams_cap));
    c->d_teams = nullptr;
    c->d_team_first = nullptr;
    c->teams_cap = 0;
    const int cap = std::max(nteam, 64);
    int rc = dev_alloc(c, (void **)&c->d_teams, team_bytes(cap));
    if (rc) return rc;
    c->d_team_first = (int *)(c->d_teams + cap);
    c->teams_cap = cap;
    return FDGPU_OK;
}

// Team sizes of the leading tables of a largest-first queue (sizes n, nodes)
// on `cap` workgroups.  A table's time is modelled as its node count times
// the measured time ratio of a team of K (one table of 4 M nodes alone on
// MI355X: 2 workgroups 0.71 of one workgroup's time, 4: 0.52, 8: 0.41, 16:
// 0.37).  For a batch time T each table gets the smallest team that finishes
// within T.  load > 0: the smallest T whose workgroup time fits `cap`
// workgroups at that load (1: every workgroup busy for T).  load 0: the T
// (among 64 between the largest table's time on 16 workgroups and on one)
// whose plan has the shortest makespan when its tickets are handed out in
// queue order, as the kernel does (a team starts once K workgroups are
// free).  Returns the team sizes of the leading tables (sizes 1 end the list).
std::vector<int> plan_teams(const std::vector<long long> &n, int cap, double load) {
    static const double ratio[17] = {0, 1.0, 0.71, 0.60, 0.52, 0.48, 0.45, 0.43, 0.41,
                                     0.40, 0.395, 0.39, 0.385, 0.38, 0.375, 0.372, 0.37};
    std::vector<int> K(n.size(), 1);
    if (n.empty() || cap < 2) return {};
    auto assign = [&](double T) {
        double cost = 0.0;
        for (size_t i = 0; i < n.size(); i++) {
            int k = 1;
            while (k < fdsweep::kTeamMax && k < cap && n[i] >= kTeamMinNodes && (double)n[i] * ratio[k] > T) k++;
            K[i] = k;
            cost += k * (double)n[i] * ratio[k];
        }
        return cost;
    };
    const double lo0 = (double)n[0] * ratio[fdsweep::kTeamMax], hi0 = (double)n[0];
    if (load > 0.0) {
        double lo = lo0, hi = hi0;
        for (int it = 0; it < 60; it++) {
            const double mid = 0.5 * (lo + hi);
            if (assign(mid) <= cap * mid * load) hi = mid;
            else lo = mid;
        }
        assign(hi);
    } else {
        // list schedule of the tickets: workgroups' free times in a min-heap
        auto makespan = [&]() {
            std::priority_queue<double, std::vector<double>, std::greater<double>> fr;
            for (int w = 0; w < cap; w++) fr.push(0.0);
            double end = 0.0;
            std::vector<double> got;
            for (size_t i = 0; i < n.size(); i++) {
                got.clear();
                for (int k = 0; k < K[i]; k++) {
                    got.push_back(fr.top());
                    fr.pop();
                }
                const double t = *std::max_element(got.begin(), got.end()) + (double)n[i] * ratio[K[i]];
                for (int k = 0; k < K[i]; k++) fr.push(t);
                end = std::max(end, t);
            }
            return end;
        };
        double bestT = hi0, best = -1.0;
        for (int j = 0; j <= 64; j++) {
            const double T = lo0 + (hi0 - lo0) * j / 64.0;
            assign(T);
            const double m = makespan();
            if (best < 0.0 || m < best * (1.0 - 1e-9)) {
                best = m;
                bestT = T;
            }
        }
        assign(bestT);
    }
    size_t m = 0;
    while (m < K.size() && K[m] > 1) m++;
    K.resize(m);
    return K;
}

// The batched launch as a graph (a queue-reset kernel, then the persistent
// sweep kernel), captured once per (build, table count, workgroups) and
// replayed; descriptors and scratch pointers are fixed for its lifetime.
int sweep_graph(fdgpu_ctx *c, int variant, int n, int nslots, int nteam, hipGraphExec_t *out) {
    // n < 2^24, nslots < 2^20, nteam < 2^11 (the team sizes are read from
    // d_team_first at run time)
    const long long key = ((long long)variant << 60) | ((long long)nteam << 44) | ((long long)nslots << 24) |
                          (long long)n;
    auto it = c->graphs.find(key);
    if (it != c->graphs.end()) { *out = it->second; return FDGPU_OK; }
    hipGraph_t graph = nullptr;
    FD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = fd_launch_reset(c->d_queue, c->d_teams, nteam, c->stream);
    hipError_t e2 = fd_launch_sweep(variant, c->d_jobs, n, c->d_slots, nslots, c->d_queue, c->d_teams,
                                    c->d_team_first, nteam, c->stream);
    hipError_t e3 = hipStreamEndCapture(c->stream, &graph);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return FDGPU_EHIP;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t e4 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e4 != hipSuccess) return FDGPU_EHIP;
    c->graphs[key] = exec;
    *out = exec;
    return FDGPU_OK;
}

int prepare_table(const fdgpu_table_desc &d, int mnx, int mny, int mnz, Table &T, std::vector<float> &preset) {
    const fdgpu_par &p = d.par;
    if (p.nx < 1 || p.ny < 1 || p.nz < 1 || d.ox < 0 || d.oy < 0 || d.oz < 0 || d.ox + p.nx > mnx ||
        d.oy + p.ny > mny || d.oz + p.nz > mnz || p.ncube < 0)
        return FDGPU_EINPUT;
    T.par = p;
    T.model = d.model;
    T.ox = d.ox; T.oy = d.oy; T.oz = d.oz;
    T.n = (long long)p.nx * p.ny * p.nz;
    const fdprep::ParValues v = to_values(p);
    int rc = fdprep::prepare_geometry(v, T.prep);
    if (rc) return rc;
    if (p.srctype != 1) {
        if (!d.preset) return FDGPU_EINPUT;
        preset.resize((size_t)T.n);
    }
    if (p.fill == 1 && !d.boxvals) return FDGPU_EINPUT;
    return fdprep::prepare_source(v, T.prep, d.preset, d.boxvals, p.srctype != 1 ? preset.data() : nullptr);
}

void release_table(fdgpu_ctx *c, Table &t) {
    dev_free(c, t.d_time, sizeof(float) * (size_t)t.n);
    dev_free(c, t.d_aux, t.bytes);
    t.d_time = nullptr;
    t.d_aux = nullptr;
}

// Solves tables whose device buffers and descriptors exist (T.job): the
// per-solve fields (reverse-pass skip bound from the model's current
// velocity range, schedule), slots and descriptors sized, one captured
// persistent launch (FDGPU_BATCHED) or one launch per table (FDGPU_PLAIN),
// then each table's counters.  Allocates only when the slots or the
// descriptor array must grow past every earlier launch of the context.
int run_tables(fdgpu_ctx *c, const std::vector<Table *> &tabs, int mode, bool no_wave) {
    const int n = (int)tabs.size();
    if (n == 0) return FDGPU_OK;
    long long maxF = 1, maxN = 1, geo_need = 0;
    int maxD = 1;
    std::vector<fdsweep::TableJob> jobs((size_t)n);
    for (int t = 0; t < n; t++) {
        Table &T = *tabs[(size_t)t];
        const Model &m = c->models[T.model];
        maxF = std::max(maxF, fdprep::max_face(T.par.nx, T.par.ny, T.par.nz));
        maxN = std::max(maxN, T.n);
        maxD = std::max(maxD, T.job.maxdim);
        geo_need = std::max(geo_need, fd_geo_lds_need(T.par.ny, T.par.nz));
        fdsweep::TableJob &J = jobs[(size_t)t];
        J = T.job;
        J.skip_ok = fdprep::skip_safe(to_values(T.par), T.prep, m.vmin, m.vmax) ? 1 : 0;
        J.change_budget = c->change_budget;
        // wavefront passes for tables the kernel initialises itself (a failed
        // wavefront is solved again face by face, which needs no host preset)
        J.wave = (J.init_mode != 1 && !no_wave) ? (c->wave & 3) : 0;
    }
    long long total_nodes = 0;
    for (int t = 0; t < n; t++) total_nodes += tabs[(size_t)t]->n;
    // queue order: largest table first (longest-processing-time-first)
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    if (mode == FDGPU_BATCHED)
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tabs[(size_t)a]->n > tabs[(size_t)b]->n; });
    std::vector<fdsweep::TableJob> qjobs((size_t)n);
    for (int t = 0; t < n; t++) qjobs[(size_t)t] = jobs[(size_t)order[(size_t)t]];
    // wide workgroups while every table gets a CU of its own, else narrow ones
    int variant = FD_WIDE;
    if (mode == FDGPU_BATCHED) {
        if (c->lanes == kNarrowThreads) variant = FD_NARROW;
        else if (c->lanes == 0 && n > c->num_cu * c->wg_per_cu[FD_WIDE]) variant = FD_NARROW;
        if (geo_need > fd_geo_lds_capacity(variant)) variant = FD_WIDE;   // same results, larger LDS
    }
    const int cap = c->num_cu * c->wg_per_cu[variant];
    // teams: the largest tables' wavefront passes shared by several
    // workgroups each -- planned (plan_teams) or, with a fixed size, the
    // leading tables of at least team_share % of the mean nodes per workgroup
    // of a full device (at most as many as the idle workgroups hold when the
    // tables are fewer, else a size-th of the workgroups) -- for tables whose
    // passes can run as wavefronts
    std::vector<int> teamK;
    if (mode == FDGPU_BATCHED && !no_wave && c->team != 1 && variant == FD_WIDE) {
        const int arena = variant == FD_NARROW ? kNarrowArenaWords : fdsweep::kArenaWords;
        int lead = 0;   // leading tables that can run as a team
        std::vector<long long> sizes((size_t)n);
        for (int t = 0; t < n; t++) {
            const fdsweep::TableJob &J = qjobs[(size_t)t];
            sizes[(size_t)t] = (long long)J.nx * J.ny * J.nz;
            if (lead == t && (J.wave & 1) && sizes[(size_t)t] < (1LL << 29) &&
                fdsweep::wave_fits(J.nx, J.ny, J.nz, arena))
                lead++;
        }
        if (c->team == 0) {
            teamK = plan_teams(sizes, cap, c->team_share / 100.0);
        } else {
            const int K = std::min(fdsweep::kTeamMax, c->team);
            const double share = (double)total_nodes / cap * (c->team_share > 0 ? c->team_share : 200) / 100.0;
            const int most = n < cap ? (cap - n) / (K - 1) : cap / K;
            for (int t = 0; t < n && (int)teamK.size() < most && (double)sizes[(size_t)t] >= share &&
                            sizes[(size_t)t] >= kTeamMinNodes;
                 t++)
                teamK.push_back(K);
        }
        if ((int)teamK.size() > lead) teamK.resize((size_t)lead);
        if (teamK.size() > 2047) teamK.resize(2047);
    }
    int nteam = (int)teamK.size(), extra = 0;
    for (int k : teamK) extra += k - 1;
    int nslots = mode == FDGPU_BATCHED ? std::min(n + extra, cap) : 1;
    if (nslots < 1) nslots = 1;
    int rc = ensure_scratch(c, nslots, maxF, maxN, maxD, !no_wave && c->wave != 0);
    if (nteam > 0 && (rc == FDGPU_ENOMEM || !c->slot_wave)) {
        // the team members' extra slots do not fit the budget (or the slots
        // hold no wavefront state): one workgroup per table
        nteam = 0;
        nslots = std::max(1, std::min(n, cap));
        rc = ensure_scratch(c, nslots, maxF, maxN, maxD, !no_wave && c->wave != 0);
    }
    if (rc) return rc;
    rc = ensure_jobs(c, n);
    if (rc) return rc;
    // the slots must hold the team tables' wavefront state (the leader's slot
    // serves the team)
    if (nteam > 0 && (!c->slot_wave || qjobs[0].nx * (long long)qjobs[0].ny * qjobs[0].nz > c->slot_node_cap))
        nteam = 0;
    std::vector<int> first((size_t)nteam + 1, 0);
    if (nteam > 0) {
        rc = ensure_teams(c, nteam);
        if (rc) return rc;
        for (int i = 0; i < nteam; i++) first[(size_t)i + 1] = first[(size_t)i] + teamK[(size_t)i];
    }
    fdsweep::TableJob *d_jobs = c->d_jobs;
    bool ok = hipMemcpyAsync(d_jobs, qjobs.data(), sizeof(fdsweep::TableJob) * (size_t)n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess;
    if (nteam > 0)
        ok = ok && hipMemcpyAsync(c->d_team_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice,
                                  c->stream) == hipSuccess;
    ok = ok && hipEventRecord(c->ev0, c->stream) == hipSuccess;
    if (mode == FDGPU_BATCHED) {
        // one captured graph per launch shape: queue reset + persistent sweep
        hipGraphExec_t g = nullptr;
        ok = ok && sweep_graph(c, variant, n, nslots, nteam, &g) == FDGPU_OK;
        ok = ok && hipGraphLaunch(g, c->stream) == hipSuccess;
    } else {
        for (int t = 0; t < n && ok; t++) {
            ok = fd_launch_reset(c->d_queue, nullptr, 0, c->stream) == hipSuccess;
            ok = ok && fd_launch_sweep(FD_WIDE, d_jobs + t, 1, c->d_slots, 1, c->d_queue, nullptr, nullptr, 0,
                                       c->stream) == hipSuccess;
        }
    }
    ok = ok && hipEventRecord(c->ev1, c->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) return FDGPU_EHIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->last_ms = ms;
    int kerr = 0;
    std::vector<Table *> again;
    for (int t = 0; t < n; t++) {
        Table &T = *tabs[(size_t)t];
        T.stats.assign(fdsweep::kStatCount, 0);
        FD_HIP(hipMemcpy(T.stats.data(), T.d_stats, sizeof(int) * fdsweep::kStatCount, hipMemcpyDeviceToHost));
        // bit 7: a wavefront pass could not finish (list capacity, no progress);
        // bits 2-5: one of the kernel's consistency checks failed (a LEVEL
        // face left nodes unprocessed or ran out of rounds, a CHANGE face ran
        // out of rounds, a skipped node did not reproduce its value).  The
        // table is solved again face by face, once; a second failure is an
        // error.  bit0 (NaN keys) is reported, not fatal.  Tables whose
        // initial grid the host supplied (srctype 2/3) cannot run again.
        if (!no_wave && (T.stats[fdsweep::kStatError] & ~1) && T.job.init_mode != 1) {
            again.push_back(&T);
            if (T.stats[fdsweep::kStatError] & ~(1 | 128)) c->resolved_again++;
            else c->wave_again++;
        } else if (T.stats[fdsweep::kStatError] & ~1) {
            kerr = 1;
        }
    }
    if (!again.empty()) {
        const float ms0 = c->last_ms;
        // the first attempt's failure reason and team size stay readable
        std::vector<std::pair<int, int>> first_try;
        for (Table *T : again) first_try.emplace_back(T->stats[fdsweep::kStatWaveFail], T->stats[fdsweep::kStatTeam]);
        const int rc2 = run_tables(c, again, mode, true);
        for (size_t q = 0; q < again.size(); q++) {
            again[q]->stats[fdsweep::kStatWaveFail] = first_try[q].first;
            again[q]->stats[fdsweep::kStatTeam] = first_try[q].second;
        }
        c->last_ms += ms0;
        if (rc2 == FDGPU_EKERNEL) kerr = 1;
        else if (rc2) return rc2;
    }
    return kerr ? FDGPU_EKERNEL : FDGPU_OK;
}

}  // namespace fdgpu_detail

using namespace fdgpu_detail;

extern "C" {

void fdgpu_par_defaults(fdgpu_par *p) {
    memset(p, 0, sizeof *p);
    p->y0 = 90.0;
    p->iminus = p->iplus = p->jminus = p->jplus = p->kminus = p->kplus = 1;
    p->floatsrc = 1;
    p->srctype = 1;
    p->ncube = 2;
    p->reverse = 1;
    p->headpref = 6;
    p->parse = 1;
    p->maxoff = -1.f;
}

const char *fdgpu_strerror(int code) {
    switch (code) {
    case FDGPU_OK: return "ok";
    case FDGPU_EHIP: return "HIP runtime error";
    case FDGPU_ENOMEM: return "out of device memory or HBM budget";
    case FDGPU_ESOURCE: return "source does not appear to be in the model";
    case FDGPU_ESRCTYPE: return "incorrect value of srctype";
    case FDGPU_EINPUT: return "bad argument or missing input data";
    case FDGPU_EIO: return "file open/read/write failure";
    case FDGPU_EPAR: return "required parameter missing";
    case FDGPU_ESTOP: return "stop due to STOP in input";
    case FDGPU_EKERNEL: return "sweep kernel reported an internal error";
    case FDGPU_EBUSY: return "busy: serving arena, staging or queue full, or call not allowed while serving";
    case FDGPU_ETIMEOUT: return "timed out waiting for a request";
    case FDGPU_ESCRATCH: return "table too large for the fixed device scratch of fdgpu_solve_noalloc";
    case FDGPU_ESPACE: return "derivative arrays too small for the traced rays";
    }
    return "unknown error";
}

int fdgpu_open(int device, unsigned long long budget, fdgpu_ctx **out) {
    if (!out) return FDGPU_EINPUT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return FDGPU_EHIP;
    FD_HIP(hipSetDevice(device));
    fdgpu_ctx *c = new fdgpu_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return FDGPU_EHIP; }
    c->num_cu = prop.multiProcessorCount;
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) { delete c; return FDGPU_EHIP; }
    c->budget = budget ? budget : (unsigned long long)(0.9 * (double)freeb);
    for (int v = 0; v < 2; v++) {
        int per_cu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fd_sweep_kernel_ptr(v),
                                                         v == FD_NARROW ? kNarrowThreads : kWideThreads, 0) == hipSuccess &&
            per_cu > 0)
            c->wg_per_cu[v] = per_cu;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev_a) != hipSuccess || hipEventCreate(&c->ev_b) != hipSuccess ||
        hipMalloc(&c->d_queue, sizeof(int)) != hipSuccess ||
        fd_noalloc_scratch((void **)&c->d_fixed) != hipSuccess) {
        fdgpu_close(c);
        return FDGPU_EHIP;
    }
    *out = c;
    return FDGPU_OK;
}

int fdgpu_close(fdgpu_ctx *c) {
    if (!c) return FDGPU_OK;
    writer_shutdown(c);   // background writes finish before their tables go
    (void)hipSetDevice(c->device);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        (void)serve_stop_locked(c);
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &kv : c->tables) release_table(c, kv.second);
    for (auto &kv : c->models) dev_free(c, kv.second.d_slow, kv.second.bytes);
    clear_graphs(c);
    clear_geos(c);
    dev_free(c, c->d_coarse, c->coarse_bytes);
    dev_free(c, c->d_jobs, sizeof(fdsweep::TableJob) * (size_t)c->jobs_cap);
    dev_free(c, c->d_scratch, c->scratch_bytes);
    dev_free(c, c->d_slots, sizeof(fdsweep::Scratch) * (size_t)c->slots_cap);
    dev_free(c, c->ray_cols, c->ray_cols_bytes);
    dev_free(c, c->ray_blk, c->ray_blk_bytes);
    dev_free(c, c->ray_out, c->ray_out_bytes);
    dev_free(c, c->d_teams, team_bytes(c->teams_cap));
    if (c->d_queue) (void)hipFree(c->d_queue);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_a) (void)hipEventDestroy(c->ev_a);
    if (c->ev_b) (void)hipEventDestroy(c->ev_b);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FDGPU_OK;
}

int fdgpu_model_create(fdgpu_ctx *c, int nx, int ny, int nz, const float *vel, int *model) {
    if (!c || !vel || !model || nx <= 0 || ny <= 0 || nz <= 0) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;   // no allocation while the request loop runs
    FD_HIP(hipSetDevice(c->device));
    Model m;
    m.nx = nx; m.ny = ny; m.nz = nz;
    const long long n = (long long)nx * ny * nz;
    m.bytes = sizeof(float) * (size_t)n;
    int rc = dev_alloc_resident(c, (void **)&m.d_slow, m.bytes);
    if (rc) return rc;
    const int id = c->next_id++;
    c->models[id] = m;
    rc = fdgpu_model_update_locked(c, id, vel);
    if (rc) {   // not registered: its memory goes with it
        dev_free(c, m.d_slow, m.bytes);
        c->models.erase(id);
        return rc;
    }
    *model = id;
    return FDGPU_OK;
}

int fdgpu_model_update(fdgpu_ctx *c, int model, const float *vel) {
    if (!c || !vel) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;
    FD_HIP(hipSetDevice(c->device));
    return fdgpu_model_update_locked(c, model, vel);
}

int fdgpu_model_release(fdgpu_ctx *c, int model) {
    if (!c) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;
    auto it = c->models.find(model);
    if (it == c->models.end()) return FDGPU_EINPUT;
    (void)hipStreamSynchronize(c->stream);
    dev_free(c, it->second.d_slow, it->second.bytes);
    c->models.erase(it);
    return FDGPU_OK;
}

int fdgpu_solve(fdgpu_ctx *c, const fdgpu_table_desc *desc, int n, int mode, int *ids) {
    if (!c || (!desc && n > 0) || n < 0 || !ids) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;
    FD_HIP(hipSetDevice(c->device));
    // prepare every table on the host first: no device work on bad input
    std::vector<Table> tabs((size_t)n);
    std::vector<std::vector<float>> presets((size_t)n);
    for (int t = 0; t < n; t++) {
        const fdgpu_table_desc &d = desc[t];
        auto mit = c->models.find(d.model);
        if (mit == c->models.end()) return FDGPU_EINPUT;
        const Model &m = mit->second;
        int rc = prepare_table(d, m.nx, m.ny, m.nz, tabs[(size_t)t], presets[(size_t)t]);
        if (rc) return rc;
        // every table's geometry row/layer tables must fit the LDS of the wide build
        if (fd_geo_lds_need(d.par.ny, d.par.nz) > fd_geo_lds_capacity(FD_WIDE)) return FDGPU_EINPUT;
    }
    // device buffers per table: the time grid and the host-prepared arrays
    std::map<std::vector<unsigned long long>, int> geo_keep;
    int rc = FDGPU_OK;
    int made = 0;
    for (int t = 0; t < n && rc == FDGPU_OK; t++) {
        Table &T = tabs[(size_t)t];
        const Model &m = c->models[T.model];
        const size_t b_sin = align256(sizeof(double) * T.prep.sinq.size());
        const size_t b_idx = align256(sizeof(int) * T.prep.init_idx.size());
        const size_t b_dist = align256(sizeof(double) * T.prep.init_dist.size());
        const size_t b_val = align256(sizeof(float) * T.prep.init_val.size());
        const size_t b_st = align256(sizeof(int) * fdsweep::kStatCount);
        T.bytes = b_sin + b_idx + b_dist + b_val + b_st;
        rc = dev_alloc_resident(c, (void **)&T.d_time, sizeof(float) * (size_t)T.n);
        if (rc) break;
        rc = dev_alloc_resident(c, (void **)&T.d_aux, T.bytes);
        if (rc) { dev_free(c, T.d_time, sizeof(float) * (size_t)T.n); T.d_time = nullptr; break; }
        made++;
        std::vector<char> aux(T.bytes, 0);
        size_t o = 0;
        memcpy(aux.data() + o, T.prep.sinq.data(), sizeof(double) * T.prep.sinq.size());
        const double *d_sin = (const double *)(T.d_aux + o);
        o += b_sin;
        memcpy(aux.data() + o, T.prep.init_idx.data(), sizeof(int) * T.prep.init_idx.size());
        const int *d_idx = (const int *)(T.d_aux + o);
        o += b_idx;
        memcpy(aux.data() + o, T.prep.init_dist.data(), sizeof(double) * T.prep.init_dist.size());
        const double *d_dist = (const double *)(T.d_aux + o);
        o += b_dist;
        memcpy(aux.data() + o, T.prep.init_val.data(), sizeof(float) * T.prep.init_val.size());
        const float *d_val = (const float *)(T.d_aux + o);
        o += b_val;
        T.d_stats = (int *)(T.d_aux + o);
        if (hipMemcpyAsync(T.d_aux, aux.data(), T.bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            rc = FDGPU_EHIP;
            break;
        }
        if (T.par.srctype != 1 &&
            hipMemcpyAsync(T.d_time, presets[(size_t)t].data(), sizeof(float) * (size_t)T.n, hipMemcpyHostToDevice,
                           c->stream) != hipSuccess) {
            rc = FDGPU_EHIP;
            break;
        }
        (void)hipStreamSynchronize(c->stream);   // host staging buffers are freed at scope end
        fdsweep::TableJob &J = T.job;
        memset(&J, 0, sizeof J);
        J.nx = T.par.nx; J.ny = T.par.ny; J.nz = T.par.nz;
        J.h = T.prep.h; J.dq = T.prep.dq; J.df = T.prep.df; J.y0 = T.prep.y0; J.z0r = T.prep.z0r;
        J.sinq = d_sin;
        J.s_sy = m.nx;
        J.s_sz = (long long)m.nx * m.ny;
        J.slow = m.d_slow + (long long)T.oz * J.s_sz + (long long)T.oy * J.s_sy + T.ox;
        J.time = T.d_time;
        J.init_mode = T.prep.init_mode;
        J.src_sidx = (long long)T.prep.zs * J.s_sz + (long long)T.prep.ys * J.s_sy + T.prep.xs;
        J.ninit = (int)T.prep.init_idx.size();
        J.init_idx = d_idx;
        J.init_dist = d_dist;
        J.init_val = d_val;
        memcpy(J.box, T.prep.box, sizeof J.box);
        memcpy(J.dflag, T.prep.dflag, sizeof J.dflag);
        memcpy(J.grow, T.prep.grow, sizeof J.grow);
        J.reverse = T.prep.reverse; J.headpref = T.prep.headpref; J.maxrad = T.prep.maxrad;
        J.radius0 = T.prep.radius0; J.srcwall = T.prep.srcwall;
        J.stats = T.d_stats;
        J.mod_stamp = nullptr;   // the solving slot's stamps
        J.last_full = nullptr;
        J.maxdim = std::max(T.par.nx, std::max(T.par.ny, T.par.nz));
        rc = geometry_for(c, J, T.prep.sinq, geo_keep, &J.geo);
        if (rc) break;
    }
    auto unwind = [&]() {
        for (int t = 0; t < made; t++) release_table(c, tabs[(size_t)t]);
    };
    if (rc) { unwind(); return rc; }
    std::vector<Table *> run((size_t)n);
    for (int t = 0; t < n; t++) run[(size_t)t] = &tabs[(size_t)t];
    rc = run_tables(c, run, mode, false);
    if (rc && rc != FDGPU_EKERNEL) { unwind(); return rc; }
    for (int t = 0; t < n; t++) {
        const int id = c->next_id++;
        ids[t] = id;
        c->tables[id] = std::move(tabs[(size_t)t]);
    }
    return rc;
}

int fdgpu_context_counters(fdgpu_ctx *c, long long *out, int n) {
    if (!c || n < 0 || (n > 0 && !out)) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (n > 0) out[0] = c->resolved_again;
    if (n > 1) out[1] = c->wave_again;
    return FDGPU_OK;
}

int fdgpu_resolve(fdgpu_ctx *c, const int *ids, int n, int mode) {
    if (!c || (!ids && n > 0) || n < 0) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;
    FD_HIP(hipSetDevice(c->device));
    std::vector<Table *> run((size_t)n);
    for (int t = 0; t < n; t++) {
        auto it = c->tables.find(ids[t]);
        if (it == c->tables.end() || it->second.from_file || it->second.par.srctype != 1) return FDGPU_EINPUT;
        if (writer_busy(c, ids[t])) return FDGPU_EBUSY;   // a background write still reads it
        if (c->models.find(it->second.model) == c->models.end()) return FDGPU_EINPUT;
        run[(size_t)t] = &it->second;
    }
    return run_tables(c, run, mode, false);
}

int fdgpu_solve_noalloc(fdgpu_ctx *c, const fdgpu_table_desc *d, float *d_time, int *stats, int nstats) {
    if (!c || !d || !d_time || (nstats > 0 && !stats)) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;   // the request loop holds every CU on the stream
    if (!c->d_fixed) return FDGPU_EHIP;
    FD_HIP(hipSetDevice(c->device));
    auto mit = c->models.find(d->model);
    if (mit == c->models.end()) return FDGPU_EINPUT;
    const Model &m = mit->second;
    Table T;
    std::vector<float> preset;
    int rc = prepare_table(*d, m.nx, m.ny, m.nz, T, preset);
    if (rc) return rc;
    if (fd_geo_lds_need(T.par.ny, T.par.nz) > fd_geo_lds_capacity(FD_WIDE)) return FDGPU_EINPUT;
    // layout of the fixed array: the host-prepared arrays, counters, last
    // full-plane marks, geometry tables, then one workgroup's face scratch
    const long long maxF = fdprep::max_face(T.par.nx, T.par.ny, T.par.nz);
    const int maxdim = std::max(T.par.nx, std::max(T.par.ny, T.par.nz));
    const size_t b_sin = align256(sizeof(double) * T.prep.sinq.size());
    const size_t b_idx = align256(sizeof(int) * T.prep.init_idx.size());
    const size_t b_dist = align256(sizeof(double) * T.prep.init_dist.size());
    const size_t b_val = align256(sizeof(float) * T.prep.init_val.size());
    const size_t b_st = align256(sizeof(int) * fdsweep::kStatCount);
    const size_t b_lf = align256(sizeof(int) * 6 * (size_t)maxdim);
    const size_t b_geo = align256(sizeof(double) * (size_t)fdsweep::geo_doubles(T.par.ny, T.par.nz));
    const size_t b_host = b_sin + b_idx + b_dist + b_val;
    const size_t need = b_host + b_st + b_lf + b_geo + slot_bytes(maxF, 0, 0, false);
    if (need > FD_NOALLOC_SCRATCH_BYTES) return FDGPU_ESCRATCH;
    char *base = c->d_fixed;
    std::vector<char> aux(b_host, 0);
    memcpy(aux.data(), T.prep.sinq.data(), sizeof(double) * T.prep.sinq.size());
    memcpy(aux.data() + b_sin, T.prep.init_idx.data(), sizeof(int) * T.prep.init_idx.size());
    memcpy(aux.data() + b_sin + b_idx, T.prep.init_dist.data(), sizeof(double) * T.prep.init_dist.size());
    memcpy(aux.data() + b_sin + b_idx + b_dist, T.prep.init_val.data(), sizeof(float) * T.prep.init_val.size());
    FD_HIP(hipMemcpyAsync(base, aux.data(), b_host, hipMemcpyHostToDevice, c->stream));
    if (T.par.srctype != 1)
        FD_HIP(hipMemcpyAsync(d_time, preset.data(), sizeof(float) * (size_t)T.n, hipMemcpyHostToDevice, c->stream));
    fdsweep::TableJob J;
    memset(&J, 0, sizeof J);
    J.nx = T.par.nx; J.ny = T.par.ny; J.nz = T.par.nz;
    J.h = T.prep.h; J.dq = T.prep.dq; J.df = T.prep.df; J.y0 = T.prep.y0; J.z0r = T.prep.z0r;
    J.sinq = (const double *)base;
    J.s_sy = m.nx;
    J.s_sz = (long long)m.nx * m.ny;
    J.slow = m.d_slow + (long long)T.oz * J.s_sz + (long long)T.oy * J.s_sy + T.ox;
    J.time = d_time;
    J.init_mode = T.prep.init_mode;
    J.src_sidx = (long long)T.prep.zs * J.s_sz + (long long)T.prep.ys * J.s_sy + T.prep.xs;
    J.ninit = (int)T.prep.init_idx.size();
    J.init_idx = (const int *)(base + b_sin);
    J.init_dist = (const double *)(base + b_sin + b_idx);
    J.init_val = (const float *)(base + b_sin + b_idx + b_dist);
    memcpy(J.box, T.prep.box, sizeof J.box);
    memcpy(J.dflag, T.prep.dflag, sizeof J.dflag);
    memcpy(J.grow, T.prep.grow, sizeof J.grow);
    J.reverse = T.prep.reverse; J.headpref = T.prep.headpref; J.maxrad = T.prep.maxrad;
    J.radius0 = T.prep.radius0; J.srcwall = T.prep.srcwall;
    int *d_stats = (int *)(base + b_host);
    J.stats = d_stats;
    J.mod_stamp = nullptr;   // no per-node stamps: the reverse-pass skip is off (same results)
    J.skip_ok = 0;
    J.last_full = (int *)(base + b_host + b_st);
    J.maxdim = maxdim;
    J.change_budget = c->change_budget;
    J.geo_fill = (double *)(base + b_host + b_st + b_lf);
    // one slot of face scratch, laid out as fill_slots lays out each slot
    char *sb = base + b_host + b_st + b_lf + b_geo;
    const size_t f = align256(sizeof(float) * (size_t)maxF);
    fdsweep::Scratch W;
    size_t o = 0;
    auto take = [&](size_t n) { char *p = sb + o; o += n; return p; };
    W.O = (float *)take(f); W.key = (unsigned *)take(f); W.key2 = (unsigned *)take(f);
    W.val = (int *)take(f); W.val2 = (int *)take(f); W.rank = (int *)take(f); W.ord = (int *)take(f);
    W.pend = (int *)take(f); W.stat = (int *)take(f); W.tout0 = (float *)take(f); W.tout4 = (float *)take(f);
    W.list0 = (int *)take(f); W.list1 = (int *)take(f); W.list2 = (int *)take(f); W.stamp = (int *)take(f);
    W.flag = (unsigned char *)take(align256((size_t)maxF));
    W.mstamp = nullptr;
    W.lastf = nullptr;
    W.wst = nullptr; W.wt0 = nullptr; W.wt4 = nullptr;   // no wavefront state: faces one by one
    W.wcnt = nullptr; W.wtmp = nullptr; W.wready = nullptr; W.wclose = nullptr;
    W.wcap = 0; W.wn = 0; W.wround = nullptr;
    bool ok = hipEventRecord(c->ev0, c->stream) == hipSuccess;
    ok = ok && fd_launch_one(&J, &W, c->stream) == hipSuccess;
    ok = ok && hipEventRecord(c->ev1, c->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) return FDGPU_EHIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->last_ms = ms;
    int st[fdsweep::kStatCount];
    FD_HIP(hipMemcpy(st, d_stats, sizeof st, hipMemcpyDeviceToHost));
    for (int q = 0; q < nstats && q < fdsweep::kStatCount; q++) stats[q] = st[q];
    return (st[fdsweep::kStatError] & ~1) ? FDGPU_EKERNEL : FDGPU_OK;
}

double fdgpu_last_kernel_ms(fdgpu_ctx *c) { return c ? c->last_ms : 0.0; }

int fdgpu_set_schedule(fdgpu_ctx *c, int change_budget) {
    if (!c || change_budget < 0) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    c->change_budget = change_budget;
    return FDGPU_OK;
}

int fdgpu_set_wave(fdgpu_ctx *c, int flags) {
    if (!c || flags < 0 || flags > 3) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    c->wave = flags;
    return FDGPU_OK;
}

int fdgpu_set_team(fdgpu_ctx *c, int size, int share_pct) {
    if (!c || size < 0 || size > fdsweep::kTeamMax || share_pct < 0) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    c->team = size;
    c->team_share = share_pct;
    return FDGPU_OK;
}

int fdgpu_get_team(fdgpu_ctx *c, int *size, int *share_pct) {
    if (!c || !size || !share_pct) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    *size = c->team;
    *share_pct = c->team_share;
    return FDGPU_OK;
}

int fdgpu_get_wave(fdgpu_ctx *c, int *flags) {
    if (!c || !flags) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    *flags = c->wave;
    return FDGPU_OK;
}

int fdgpu_set_workgroup(fdgpu_ctx *c, int lanes) {
    if (!c || (lanes != 0 && lanes != kNarrowThreads && lanes != kWideThreads)) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    c->lanes = lanes;
    return FDGPU_OK;
}

int fdgpu_get_schedule(fdgpu_ctx *c, int *change_budget) {
    if (!c || !change_budget) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    *change_budget = c->change_budget;
    return FDGPU_OK;
}

int fdgpu_table_read(fdgpu_ctx *c, int table, float *host, unsigned long long count) {
    if (!c || !host) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->tables.find(table);
    if (it == c->tables.end() || count < (unsigned long long)it->second.n) return FDGPU_EINPUT;
    FD_HIP(hipSetDevice(c->device));
    FD_HIP(hipMemcpyAsync(host, it->second.d_time, sizeof(float) * (size_t)it->second.n, hipMemcpyDeviceToHost,
                          c->stream));
    FD_HIP(hipStreamSynchronize(c->stream));
    return FDGPU_OK;
}

int fdgpu_table_stats(fdgpu_ctx *c, int table, int *stats, int nstats) {
    if (!c || !stats) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->tables.find(table);
    if (it == c->tables.end()) return FDGPU_EINPUT;
    for (int s = 0; s < nstats && s < (int)it->second.stats.size(); s++) stats[s] = it->second.stats[(size_t)s];
    return FDGPU_OK;
}

int fdgpu_table_release(fdgpu_ctx *c, int table) {
    if (!c) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;
    auto it = c->tables.find(table);
    if (it == c->tables.end()) return FDGPU_EINPUT;
    if (writer_busy(c, table)) return FDGPU_EBUSY;   // a background write still reads it
    (void)hipSetDevice(c->device);
    release_table(c, it->second);
    c->tables.erase(it);
    return FDGPU_OK;
}

int fdgpu_table_device_ptr(fdgpu_ctx *c, int table, void **dptr) {
    if (!c || !dptr) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    auto it = c->tables.find(table);
    if (it == c->tables.end()) return FDGPU_EINPUT;
    *dptr = it->second.d_time;
    return FDGPU_OK;
}

int fdgpu_table_write(fdgpu_ctx *c, int table, const char *path, const void *vel_header) {
    if (!c || !path) return FDGPU_EINPUT;
    std::vector<float> host;
    fdgpu_par par;
    fdprep::Prepared prep;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->tables.find(table);
        if (it == c->tables.end() || it->second.from_file) return FDGPU_EINPUT;
        host.resize((size_t)it->second.n);
        FD_HIP(hipSetDevice(c->device));
        FD_HIP(hipMemcpyAsync(host.data(), it->second.d_time, sizeof(float) * host.size(), hipMemcpyDeviceToHost,
                              c->stream));
        FD_HIP(hipStreamSynchronize(c->stream));
        par = it->second.par;
        prep = it->second.prep;
    }
    // NULL: the reference's header for a headerless velocity file
    // (sphfd.c:615-656, with the velocity swab), as the drop-in derives it
    fdio::VHead head;
    if (!vel_header) {
        memset(&head, 0, sizeof head);
        head = fdio::table_header(to_values(par), prep, head, false);
        vel_header = &head;
    }
    return fdio::write_table(path, par, prep, vel_header, host.data());
}

}  // extern "C"

// model upload (called with the context lock held): the velocities are
// copied into the model's own buffer and turned into slowness there, so an
// update allocates nothing.  The velocity range bounds the reverse-pass skip
// (skip_safe); it is stored only once the device holds the new slowness, and
// cleared (no skipping) if the upload fails part-way.
int fdgpu_model_update_locked(fdgpu_ctx *c, int model, const float *vel) {
    auto it = c->models.find(model);
    if (it == c->models.end()) return FDGPU_EINPUT;
    Model &m = it->second;
    const long long nv = (long long)m.nx * m.ny * m.nz;
    double vmin = 1e300, vmax = -1e300;
    for (long long q = 0; q < nv; q++) {
        const double v = vel[q];
        if (!(v == v)) { vmin = 0.0; break; }   // NaN: never skip
        vmin = v < vmin ? v : vmin;
        vmax = v > vmax ? v : vmax;
    }
    bool ok = hipMemcpyAsync(m.d_slow, vel, m.bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess;
    ok = ok && fd_launch_slowness(m.d_slow, m.d_slow, nv, c->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) {
        m.vmin = m.vmax = 0.0;
        return FDGPU_EHIP;
    }
    m.vmin = vmin;
    m.vmax = vmax;
    return FDGPU_OK;
}
