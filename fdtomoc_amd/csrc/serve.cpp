// serve.cpp -- continuous-batching request loop of the sphfd GPU path
// (fdgpu_serve_* in include/fdtomoc_amd.h).
//
// One persistent server kernel (sphfd_serve_kernel) runs while requests are
// outstanding.  Each of its workgroups (one per CU) claims the next published
// request, solves its table exactly as a batched launch would, posts a
// completion word to pinned host memory and claims the next one, so requests
// that arrive while others are on the GPU start as soon as a workgroup frees
// up, and their work overlaps with whatever else is running.
//
// Admission is by memory: every request takes a block of a fixed HBM arena
// (time grid, stencil tables, stamps, its own slowness when it carries its own
// velocities) and a block of a pinned staging area (its host-side inputs,
// returned when it completes).  When either is full, or the ring of request
// descriptors is, submit returns FDGPU_EBUSY and the caller retries after
// releasing finished requests.  Nothing is allocated or freed on the device
// while the loop runs: hipFree waits for the whole device, which would stall
// the host until the server kernel went idle.
//
// The server kernel leaves when the host sets the stop word or when no
// workgroup has had a request for `idle` seconds; the host relaunches it when
// a request is outstanding and the previous launch has finished.
#include <hip/hip_runtime.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "../../include/fdtomoc_amd.h"
#include "fd_kernels.h"
#include "fdgpu_internal.h"
#include "sphfd_io.h"

using namespace fdgpu_detail;

namespace {

#define SV_HIP(call)                                  \
    do {                                              \
        hipError_t _e = (call);                       \
        if (_e != hipSuccess) return hip_fail(_e);    \
    } while (0)

constexpr double kIdleSeconds = 0.25;

// Per-request device block: time | sinq | stats | stamps | last_full |
// geometry tables | slowness (own velocities or a coarse model) | coarse
// slowness | coarse node coordinates | axis entries (coarse model)
struct DevLayout {
    size_t time, sinq, stats, stamp, lastf, geo, slow, cslow, cg, caxis, total;
};

// Sizes of a coarse request's extra parts (zero for other requests).
struct CoarseSizes {
    long long ncoarse = 0;   // coarse nodes of one phase
    int ncg = 0;             // coarse node coordinates
    long long nax = 0;       // axis entries
};

CoarseSizes coarse_sizes(const fdgpu_coarse &co, const fdgpu_par &p) {
    CoarseSizes cs;
    cs.ncoarse = (long long)co.nxc * co.nyc * co.nzc;
    cs.ncg = co.nxc + co.nyc + co.nzc;
    cs.nax = (long long)p.nx + p.ny + p.nz;
    return cs;
}

DevLayout dev_layout(const fdgpu_par &p, bool own_vel, const CoarseSizes &cs = CoarseSizes()) {
    const long long n = (long long)p.nx * p.ny * p.nz;
    const int maxdim = std::max(p.nx, std::max(p.ny, p.nz));
    DevLayout L;
    size_t o = 0;
    L.time = o; o += align256(sizeof(float) * (size_t)n);
    L.sinq = o; o += align256(sizeof(double) * (size_t)(p.ny + 2));
    L.stats = o; o += align256(sizeof(int) * fdsweep::kStatCount);
    L.stamp = o; o += align256(sizeof(int) * (size_t)n);
    L.lastf = o; o += align256(sizeof(int) * 6 * (size_t)maxdim);
    L.geo = o; o += align256(sizeof(double) * (size_t)fdsweep::geo_doubles(p.ny, p.nz));
    L.slow = o; if (own_vel) o += align256(sizeof(float) * (size_t)n);
    L.cslow = o; o += cs.ncoarse ? align256(sizeof(float) * (size_t)cs.ncoarse) : 0;
    L.cg = o; o += cs.ncg ? align256(sizeof(float) * (size_t)cs.ncg) : 0;
    L.caxis = o; o += cs.nax ? align256(sizeof(C2fAxis) * (size_t)cs.nax) : 0;
    L.total = o;
    return L;
}

// Per-request staging block: sinq | init_idx | init_dist | init_val | velocity
// | preset | coarse velocities | coarse node coordinates
struct StageLayout {
    size_t sinq, idx, dist, val, vel, preset, cvel, cg, total;
};

StageLayout stage_layout(const fdprep::Prepared &g, long long n, bool own_vel, bool preset,
                         const CoarseSizes &cs = CoarseSizes()) {
    StageLayout L;
    size_t o = 0;
    L.sinq = o; o += align256(sizeof(double) * g.sinq.size());
    L.idx = o; o += align256(sizeof(int) * g.init_idx.size());
    L.dist = o; o += align256(sizeof(double) * g.init_dist.size());
    L.val = o; o += align256(sizeof(float) * g.init_val.size());
    L.vel = o; if (own_vel) o += align256(sizeof(float) * (size_t)n);
    L.preset = o; if (preset) o += align256(sizeof(float) * (size_t)n);
    L.cvel = o; o += cs.ncoarse ? align256(sizeof(float) * (size_t)cs.ncoarse) : 0;
    L.cg = o; o += cs.ncg ? align256(sizeof(float) * (size_t)cs.ncg) : 0;
    L.total = o;
    return L;
}

void free_serve(fdgpu_ctx *c) {
    Serve &S = c->serve;
    if (S.arena) dev_free(c, S.arena, S.arena_alloc.total());
    if (S.d_scratch) dev_free(c, S.d_scratch, slot_bytes(S.max_face) * (size_t)S.nslots);
    if (S.d_slots) dev_free(c, S.d_slots, sizeof(fdsweep::Scratch) * (size_t)S.nslots);
    if (S.d_dev) dev_free(c, S.d_dev, sizeof(fdsweep::ServeDev));
    if (S.ctl) (void)hipHostFree(S.ctl);
    if (S.ring) (void)hipHostFree(S.ring);
    if (S.done) (void)hipHostFree(S.done);
    if (S.stage) (void)hipHostFree(S.stage);
    if (S.bounce) (void)hipHostFree(S.bounce);
    if (S.kdone) (void)hipEventDestroy(S.kdone);
    if (S.sstream) (void)hipStreamDestroy(S.sstream);
    if (S.cstream) (void)hipStreamDestroy(S.cstream);
    S = Serve();
}

// Picks up completion words posted by the server kernel.
int scan_done(fdgpu_ctx *c) {
    Serve &S = c->serve;
    for (auto it = S.pending.begin(); it != S.pending.end();) {
        const unsigned long long seq = it->first;
        const unsigned long long w = __atomic_load_n(&S.done[seq % (unsigned long long)S.cap], __ATOMIC_ACQUIRE);
        if (w != seq + 1) { ++it; continue; }
        ServeReq &R = S.reqs[it->second];
        R.stats.assign(fdsweep::kStatCount, 0);
        int *hb = (int *)S.bounce;
        SV_HIP(hipMemcpyAsync(hb, R.d_stats, sizeof(int) * fdsweep::kStatCount, hipMemcpyDeviceToHost, S.cstream));
        SV_HIP(hipStreamSynchronize(S.cstream));
        memcpy(R.stats.data(), hb, sizeof(int) * fdsweep::kStatCount);
        S.stage_alloc.give(R.s_off, R.s_bytes);
        R.s_bytes = 0;
        R.done = true;
        it = S.pending.erase(it);
    }
    return FDGPU_OK;
}

// Makes sure a server kernel runs while requests are outstanding.
int ensure_running(fdgpu_ctx *c) {
    Serve &S = c->serve;
    if (S.launched) {
        const hipError_t q = hipEventQuery(S.kdone);
        if (q == hipErrorNotReady) return FDGPU_OK;
        if (q != hipSuccess) return hip_fail(q);
        S.launched = false;
    }
    int rc = scan_done(c);
    if (rc) return rc;
    if (S.pending.empty()) return FDGPU_OK;
    SV_HIP(fd_launch_serve(S.ctl, S.ring, S.done, S.cap, S.d_slots, S.nslots, S.d_dev, S.idle_ticks, S.sstream));
    SV_HIP(hipEventRecord(S.kdone, S.sstream));
    S.launched = true;
    S.launches++;
    return FDGPU_OK;
}

ServeReq *find_req(fdgpu_ctx *c, int id) {
    auto it = c->serve.reqs.find(id);
    return it == c->serve.reqs.end() ? nullptr : &it->second;
}

}  // namespace

namespace fdgpu_detail {

int serve_stop_locked(fdgpu_ctx *c) {
    Serve &S = c->serve;
    if (!S.on) return FDGPU_OK;
    (void)hipSetDevice(c->device);
    // drain: with the stop word set, workgroups still claim every published
    // request before they leave, so each launch empties the ring
    __atomic_store_n(&S.ctl->stop, 1, __ATOMIC_RELEASE);
    int rc = FDGPU_OK;
    while (true) {
        rc = ensure_running(c);
        if (rc || !S.launched) break;
        if (hipStreamSynchronize(S.sstream) != hipSuccess) { rc = FDGPU_EHIP; break; }
        S.launched = false;
        rc = scan_done(c);
        if (rc || S.pending.empty()) break;
    }
    free_serve(c);
    return rc;
}

}  // namespace fdgpu_detail

extern "C" {

void fdgpu_serve_config_defaults(fdgpu_serve_config *cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof *cfg);
    cfg->staging_bytes = 1ull << 30;
    cfg->queue_capacity = 4096;
}

int fdgpu_serve_start(fdgpu_ctx *c, const fdgpu_serve_config *cfg) {
    if (!c || !cfg || cfg->max_nx < 1 || cfg->max_ny < 1 || cfg->max_nz < 1 || cfg->queue_capacity < 0 ||
        cfg->workgroups < 0)
        return FDGPU_EINPUT;
    // the server is the wide build: its LDS holds the geometry tables of grids up to this size
    if (fd_geo_lds_need(cfg->max_ny, cfg->max_nz) > fd_geo_lds_capacity(FD_WIDE)) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    Serve &S = c->serve;
    if (S.on) return FDGPU_EBUSY;
    SV_HIP(hipSetDevice(c->device));
    SV_HIP(hipStreamSynchronize(c->stream));
    S = Serve();
    S.maxdim[0] = cfg->max_nx; S.maxdim[1] = cfg->max_ny; S.maxdim[2] = cfg->max_nz;
    S.max_face = fdprep::max_face(cfg->max_nx, cfg->max_ny, cfg->max_nz);
    const int full = c->num_cu * c->wg_per_cu[FD_WIDE];
    S.nslots = cfg->workgroups ? std::min(cfg->workgroups, full) : full;
    S.cap = cfg->queue_capacity ? cfg->queue_capacity : 4096;
    const size_t stage_bytes = cfg->staging_bytes ? (size_t)cfg->staging_bytes : (size_t)1 << 30;
    const long long maxN = (long long)cfg->max_nx * cfg->max_ny * cfg->max_nz;
    S.bounce_bytes = std::max(sizeof(float) * (size_t)maxN, sizeof(int) * (size_t)fdsweep::kStatCount);
    int rc = FDGPU_OK;
    auto fail = [&](int code) { free_serve(c); return code; };
    // device: scratch slots, counters, then the arena from what the budget leaves
    rc = dev_alloc(c, (void **)&S.d_scratch, slot_bytes(S.max_face) * (size_t)S.nslots);
    if (rc) return fail(rc);
    rc = dev_alloc(c, (void **)&S.d_slots, sizeof(fdsweep::Scratch) * (size_t)S.nslots);
    if (rc) return fail(rc);
    rc = dev_alloc(c, (void **)&S.d_dev, sizeof(fdsweep::ServeDev));
    if (rc) return fail(rc);
    unsigned long long arena = cfg->arena_bytes;
    if (!arena) arena = c->budget > c->used ? c->budget - c->used : 0;
    arena &= ~255ull;
    if (arena < 256) return fail(FDGPU_ENOMEM);
    rc = dev_alloc(c, (void **)&S.arena, (size_t)arena);
    if (rc) return fail(rc);
    S.arena_alloc.reset((size_t)arena);
    // pinned host memory: control words, ring, completion words, staging (coherent:
    // the server kernel reads them while it runs), read-back buffer
    if (hipHostMalloc((void **)&S.ctl, sizeof(fdsweep::ServeCtl), hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&S.ring, sizeof(fdsweep::ServeJob) * (size_t)S.cap, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&S.done, sizeof(unsigned long long) * (size_t)S.cap, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&S.stage, stage_bytes, hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc((void **)&S.bounce, S.bounce_bytes, hipHostMallocDefault) != hipSuccess)
        return fail(FDGPU_ENOMEM);
    memset(S.ctl, 0, sizeof(fdsweep::ServeCtl));
    memset(S.ring, 0, sizeof(fdsweep::ServeJob) * (size_t)S.cap);
    memset(S.done, 0, sizeof(unsigned long long) * (size_t)S.cap);
    S.stage_alloc.reset(stage_bytes);
    if (hipStreamCreateWithFlags(&S.sstream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&S.cstream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&S.kdone, hipEventDisableTiming) != hipSuccess)
        return fail(FDGPU_EHIP);
    rc = fill_slots(S.d_scratch, S.nslots, S.max_face, S.d_slots);
    if (rc) return fail(rc);
    const fdsweep::ServeDev zero{};
    if (hipMemcpy(S.d_dev, &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) return fail(FDGPU_EHIP);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0)
        khz = 100000;
    S.idle_ticks = (long long)(kIdleSeconds * 1000.0 * khz);
    S.on = true;
    return FDGPU_OK;
}

namespace {

// One request: a window of a resident model, or the request's own velocities,
// or one phase of a coarse model interpolated on the device.
int submit_impl(fdgpu_ctx *c, const fdgpu_table_desc *d, const float *velocity, const fdgpu_coarse *coarse,
                int phase, int *request) {
    if (!c || !d || !request) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    Serve &S = c->serve;
    if (!S.on) return FDGPU_EINPUT;
    SV_HIP(hipSetDevice(c->device));
    const fdgpu_par &p = d->par;
    const Model *m = nullptr;
    int mnx = p.nx, mny = p.ny, mnz = p.nz;
    CoarseSizes cs;
    if (coarse) {
        int cx, cy, cz;
        if (coarse_check(coarse, cx, cy, cz) || cx != p.nx || cy != p.ny || cz != p.nz) return FDGPU_EINPUT;
        if (phase != 0 && phase != 1) return FDGPU_EINPUT;
        cs = coarse_sizes(*coarse, p);
    }
    if (velocity || coarse) {
        if (d->ox || d->oy || d->oz) return FDGPU_EINPUT;
    } else {
        auto it = c->models.find(d->model);
        if (it == c->models.end()) return FDGPU_EINPUT;
        m = &it->second;
        mnx = m->nx; mny = m->ny; mnz = m->nz;
    }
    if (fdprep::max_face(p.nx, p.ny, p.nz) > S.max_face) return FDGPU_EINPUT;   // beyond the configured size
    Table T;
    std::vector<float> preset;
    int rc = prepare_table(*d, mnx, mny, mnz, T, preset);
    if (rc) return rc;
    if ((long long)S.pending.size() >= S.cap) {
        (void)ensure_running(c);
        if ((long long)S.pending.size() >= S.cap) return FDGPU_EBUSY;
    }
    const bool own = velocity != nullptr || coarse != nullptr;   // the request brings its own slowness
    const DevLayout DL = dev_layout(p, own, cs);
    const StageLayout SL = stage_layout(T.prep, T.n, velocity != nullptr, p.srctype != 1, cs);
    const size_t a_off = S.arena_alloc.take(DL.total);
    if (a_off == SIZE_MAX) return FDGPU_EBUSY;
    const size_t s_off = S.stage_alloc.take(SL.total);
    if (s_off == SIZE_MAX) {
        S.arena_alloc.give(a_off, DL.total);
        return FDGPU_EBUSY;
    }
    // host-side inputs into the staging block
    char *st = S.stage + s_off;
    memcpy(st + SL.sinq, T.prep.sinq.data(), sizeof(double) * T.prep.sinq.size());
    memcpy(st + SL.idx, T.prep.init_idx.data(), sizeof(int) * T.prep.init_idx.size());
    memcpy(st + SL.dist, T.prep.init_dist.data(), sizeof(double) * T.prep.init_dist.size());
    memcpy(st + SL.val, T.prep.init_val.data(), sizeof(float) * T.prep.init_val.size());
    double vmin = 0.0, vmax = 0.0;   // a coarse request has no range here: no skip (results identical)
    if (coarse) {
        std::vector<float> g;
        coarse_coords(coarse, g);
        memcpy(st + SL.cvel, coarse->velocity + phase * cs.ncoarse, sizeof(float) * (size_t)cs.ncoarse);
        memcpy(st + SL.cg, g.data(), sizeof(float) * g.size());
    } else if (own) {
        memcpy(st + SL.vel, velocity, sizeof(float) * (size_t)T.n);
        vmin = 1e300; vmax = -1e300;
        for (long long q = 0; q < T.n; q++) {
            const double v = velocity[q];
            if (!(v == v)) { vmin = 0.0; break; }
            vmin = v < vmin ? v : vmin;
            vmax = v > vmax ? v : vmax;
        }
    } else {
        vmin = m->vmin;
        vmax = m->vmax;
    }
    if (p.srctype != 1) memcpy(st + SL.preset, preset.data(), sizeof(float) * (size_t)T.n);
    // the descriptor
    char *blk = S.arena + a_off;
    const unsigned long long seq = S.tail;
    fdsweep::ServeJob &E = S.ring[seq % (unsigned long long)S.cap];
    fdsweep::ServeJob job;
    memset(&job, 0, sizeof job);
    fdsweep::TableJob &J = job.job;
    J.nx = p.nx; J.ny = p.ny; J.nz = p.nz;
    J.h = T.prep.h; J.dq = T.prep.dq; J.df = T.prep.df; J.y0 = T.prep.y0; J.z0r = T.prep.z0r;
    J.sinq = (const double *)(blk + DL.sinq);
    if (own) {
        J.s_sy = p.nx;
        J.s_sz = (long long)p.nx * p.ny;
        J.slow = (const float *)(blk + DL.slow);
    } else {
        J.s_sy = m->nx;
        J.s_sz = (long long)m->nx * m->ny;
        J.slow = m->d_slow + (long long)d->oz * J.s_sz + (long long)d->oy * J.s_sy + d->ox;
    }
    J.time = (float *)(blk + DL.time);
    J.init_mode = T.prep.init_mode;
    J.src_sidx = (long long)T.prep.zs * J.s_sz + (long long)T.prep.ys * J.s_sy + T.prep.xs;
    J.ninit = (int)T.prep.init_idx.size();
    J.init_idx = (const int *)(st + SL.idx);
    J.init_dist = (const double *)(st + SL.dist);
    J.init_val = (const float *)(st + SL.val);
    memcpy(J.box, T.prep.box, sizeof J.box);
    memcpy(J.dflag, T.prep.dflag, sizeof J.dflag);
    memcpy(J.grow, T.prep.grow, sizeof J.grow);
    J.reverse = T.prep.reverse; J.headpref = T.prep.headpref; J.maxrad = T.prep.maxrad;
    J.radius0 = T.prep.radius0; J.srcwall = T.prep.srcwall;
    J.stats = (int *)(blk + DL.stats);
    J.mod_stamp = (int *)(blk + DL.stamp);
    J.last_full = (int *)(blk + DL.lastf);
    J.maxdim = std::max(p.nx, std::max(p.ny, p.nz));
    J.skip_ok = fdprep::skip_safe(to_values(p), T.prep, vmin, vmax) ? 1 : 0;
    J.change_budget = c->change_budget;
    J.geo_fill = (double *)(blk + DL.geo);   // filled by the workgroup that takes the request
    job.vel_src = velocity ? (const float *)(st + SL.vel) : nullptr;
    job.slow_dst = own ? (float *)(blk + DL.slow) : nullptr;
    job.nvel = velocity ? T.n : 0;
    if (coarse) {
        job.cvel_src = (const float *)(st + SL.cvel);
        job.cg_src = (const float *)(st + SL.cg);
        job.cslow_dst = (float *)(blk + DL.cslow);
        job.caxis = (C2fAxis *)(blk + DL.caxis);
        C2fGrid &g = job.cgrid;
        g.nxc = coarse->nxc; g.nyc = coarse->nyc; g.nzc = coarse->nzc;
        const float *gd = (const float *)(blk + DL.cg);
        g.gx = gd; g.gy = gd + coarse->nxc; g.gz = gd + coarse->nxc + coarse->nyc;
        g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
        g.x0 = coarse->x0; g.y0 = coarse->y0; g.z0 = coarse->z0; g.h = coarse->h;
    }
    job.sinq_src = (const double *)(st + SL.sinq);
    job.sinq_dst = (double *)(blk + DL.sinq);
    job.nsinq = (int)T.prep.sinq.size();
    job.preset_src = p.srctype != 1 ? (const float *)(st + SL.preset) : nullptr;
    memcpy(&E, &job, sizeof job);
    // publish
    __atomic_store_n(&S.ctl->tail, seq + 1, __ATOMIC_RELEASE);
    S.tail = seq + 1;
    const int id = c->next_id++;
    ServeReq &R = S.reqs[id];
    R.seq = seq;
    R.par = p;
    R.n = T.n;
    R.a_off = a_off; R.a_bytes = DL.total;
    R.s_off = s_off; R.s_bytes = SL.total;
    R.d_time = J.time;
    R.d_stats = J.stats;
    R.prep = T.prep;
    S.pending[seq] = id;
    *request = id;
    return ensure_running(c);
}

}  // namespace

int fdgpu_serve_submit(fdgpu_ctx *c, const fdgpu_table_desc *d, const float *velocity, int *request) {
    return submit_impl(c, d, velocity, nullptr, 0, request);
}

int fdgpu_serve_submit_coarse(fdgpu_ctx *c, const fdgpu_table_desc *d, const fdgpu_coarse *coarse, int phase,
                              int *request) {
    if (!coarse) return FDGPU_EINPUT;
    return submit_impl(c, d, nullptr, coarse, phase, request);
}

int fdgpu_serve_poll(fdgpu_ctx *c, int request, int *done) {
    if (!c || !done) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->serve.on) return FDGPU_EINPUT;
    ServeReq *R = find_req(c, request);
    if (!R) return FDGPU_EINPUT;
    int rc = ensure_running(c);
    *done = R->done ? 1 : 0;
    return rc;
}

int fdgpu_serve_wait(fdgpu_ctx *c, int request, double timeout_s) {
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
        int done = 0;
        const int rc = fdgpu_serve_poll(c, request, &done);
        if (rc) return rc;
        if (done) return FDGPU_OK;
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (timeout_s >= 0.0 && el > timeout_s) return FDGPU_ETIMEOUT;
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

int fdgpu_serve_read(fdgpu_ctx *c, int request, float *host, unsigned long long count) {
    if (!c || !host) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    Serve &S = c->serve;
    ServeReq *R = find_req(c, request);
    if (!S.on || !R || count < (unsigned long long)R->n) return FDGPU_EINPUT;
    if (!R->done) return FDGPU_EBUSY;
    SV_HIP(hipSetDevice(c->device));
    SV_HIP(hipMemcpyAsync(S.bounce, R->d_time, sizeof(float) * (size_t)R->n, hipMemcpyDeviceToHost, S.cstream));
    SV_HIP(hipStreamSynchronize(S.cstream));
    memcpy(host, S.bounce, sizeof(float) * (size_t)R->n);
    return FDGPU_OK;
}

int fdgpu_serve_stats(fdgpu_ctx *c, int request, int *stats, int nstats) {
    if (!c || !stats) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    ServeReq *R = find_req(c, request);
    if (!c->serve.on || !R) return FDGPU_EINPUT;
    if (!R->done) return FDGPU_EBUSY;
    for (int s = 0; s < nstats && s < (int)R->stats.size(); s++) stats[s] = R->stats[(size_t)s];
    return FDGPU_OK;
}

int fdgpu_serve_write(fdgpu_ctx *c, int request, const char *path, const void *vel_header) {
    if (!path) return FDGPU_EINPUT;
    std::vector<float> host;
    fdgpu_par par;
    fdprep::Prepared prep;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        Serve &S = c->serve;
        ServeReq *R = find_req(c, request);
        if (!S.on || !R) return FDGPU_EINPUT;
        if (!R->done) return FDGPU_EBUSY;
        host.resize((size_t)R->n);
        SV_HIP(hipSetDevice(c->device));
        SV_HIP(hipMemcpyAsync(S.bounce, R->d_time, sizeof(float) * (size_t)R->n, hipMemcpyDeviceToHost, S.cstream));
        SV_HIP(hipStreamSynchronize(S.cstream));
        memcpy(host.data(), S.bounce, sizeof(float) * (size_t)R->n);
        par = R->par;
        prep = R->prep;
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

int fdgpu_serve_request_bytes(const fdgpu_par *par, int own_velocity, const fdgpu_coarse *coarse,
                              unsigned long long *bytes) {
    if (!par || !bytes || par->nx <= 0 || par->ny <= 0 || par->nz <= 0) return FDGPU_EINPUT;
    CoarseSizes cs;
    if (coarse) {
        int cx, cy, cz;
        if (coarse_check(coarse, cx, cy, cz) || cx != par->nx || cy != par->ny || cz != par->nz) return FDGPU_EINPUT;
        cs = coarse_sizes(*coarse, *par);
    }
    *bytes = (unsigned long long)dev_layout(*par, own_velocity != 0 || coarse, cs).total;
    return FDGPU_OK;
}

int fdgpu_serve_release(fdgpu_ctx *c, int request) {
    if (!c) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    Serve &S = c->serve;
    auto it = S.reqs.find(request);
    if (!S.on || it == S.reqs.end()) return FDGPU_EINPUT;
    if (!it->second.done) return FDGPU_EBUSY;
    S.arena_alloc.give(it->second.a_off, it->second.a_bytes);
    S.reqs.erase(it);
    return FDGPU_OK;
}

int fdgpu_serve_info(fdgpu_ctx *c, unsigned long long *info, int ninfo) {
    if (!c || !info) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    const Serve &S = c->serve;
    const unsigned long long v[6] = {S.on ? 1ull : 0ull, (unsigned long long)S.pending.size(),
                                     (unsigned long long)S.arena_alloc.used(), (unsigned long long)S.arena_alloc.total(),
                                     (unsigned long long)S.launches, (unsigned long long)S.nslots};
    for (int i = 0; i < ninfo && i < 6; i++) info[i] = v[i];
    return FDGPU_OK;
}

int fdgpu_serve_stop(fdgpu_ctx *c) {
    if (!c) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    return serve_stop_locked(c);
}

}  // extern "C"
