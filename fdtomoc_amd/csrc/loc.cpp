// loc.cpp -- host side of the grid-search location (include/fdtomoc_amd.h,
// fdgpu_locate): sphfdloc's per-event loop (src/sphfdloc.c:752-1130) over
// resident travel-time tables.
//
// Each search pass runs on the device for every event still being located
// (loc_kernel.hip: the node search, then the two refinements); between
// passes the host applies the reference's outlier rules to each event's
// residuals (sphfdloc.c:1080-1126): picks beyond both limits are dropped and
// the event searched again; once per event, dropped picks the new hypocentre
// explains are restored and the event searched again.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <set>
#include <vector>

#include "fdgpu_internal.h"
#include "loc_kernels.h"

using namespace fdgpu_detail;

namespace {

constexpr int kMaxObs = 800;     // picks per event (include/common/parameter.h maxobs)
constexpr int kMaxDiv = 512;     // refinement divisions: (2*512+1)^3 points stay below 2^31

// Outlier test of one pick at the hypocentre (sphfdloc.c:1085-1087, 1115-1117).
bool within(const fdgpu_loc_params &p, float resmin, float avrmin, float tps) {
    const float absres = fabsf(resmin - avrmin);
    const float percres = 100.0f * absres / tps;
    return absres <= p.resthres || percres <= p.resthrep;
}

struct EventRun {
    int ngood = 0;
    int passes = 0;
    bool retried = false;
    bool done = false;
};

}  // namespace

void fdgpu_loc_params_defaults(fdgpu_loc_params *p) {
    if (!p) return;
    memset(p, 0, sizeof *p);
    p->kmin = 2;
    p->ndiv = 20;
    p->ndiv2 = 20;
    p->ivs = 1;
    p->vpvs = 1.78f;
    p->nthres = 8;
    p->resthres = 0.5f;
    p->resthrep = 5.0f;
    p->stdmax = 15.0f;
}

int fdgpu_locate(fdgpu_ctx *c, const fdgpu_loc_params *lp, fdgpu_loc_event *ev, int nev) {
    if (!c || !lp || (nev > 0 && !ev) || nev < 0) return FDGPU_EINPUT;
    const fdgpu_loc_params &P = *lp;
    if (P.nx < 2 || P.ny < 2 || P.nz < 2 || P.kmin < 0 || P.ndiv < 1 || P.ndiv2 < 1 || P.ndiv > kMaxDiv ||
        P.ndiv2 > kMaxDiv)
        return FDGPU_EINPUT;
    if ((long long)P.nx * P.ny * P.nz >= (1LL << 30)) return FDGPU_EINPUT;   // 32-bit byte offsets
    if (!(P.df > 0.0) || !(P.dq > 0.0) || !(P.h > 0.0)) return FDGPU_EINPUT;
    if (nev == 0) return FDGPU_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;   // no allocation while the request loop runs
    if (hipSetDevice(c->device) != hipSuccess) return FDGPU_EHIP;

    // the picks of every event, in event order, with their tables resolved
    std::vector<LocPick> picks;
    std::vector<LocEvent> events((size_t)nev);
    std::vector<EventRun> run((size_t)nev);
    for (int e = 0; e < nev; e++) {
        const fdgpu_loc_event &E = ev[e];
        if (E.npick < 1 || E.npick > kMaxObs || !E.table || !E.obstime || !E.pwt || !E.phase || !E.isgood ||
            !E.res || !E.tcalc)
            return FDGPU_EINPUT;
        events[(size_t)e].first = (int)picks.size();
        events[(size_t)e].npick = E.npick;
        for (int p = 0; p < E.npick; p++) {
            auto it = c->tables.find(E.table[p]);
            if (it == c->tables.end() || !it->second.d_time) return FDGPU_EINPUT;
            const fdgpu_par &tp = it->second.par;
            if (tp.nx != P.nx || tp.ny != P.ny || tp.nz != P.nz) return FDGPU_EINPUT;
            if (E.phase[p] != 'P' && E.phase[p] != 'S') return FDGPU_EINPUT;
            LocPick k;
            k.table = it->second.d_time;
            k.obs = E.obstime[p];
            k.pwt = E.pwt[p];
            k.good = E.isgood[p] ? 1 : 0;
            k.scale = P.ivs == 0 && E.phase[p] == 'S';
            picks.push_back(k);
        }
        // every pick counts in the first pass's variance, flagged or not (sphfdloc.c:683)
        run[(size_t)e].ngood = E.npick;
    }
    const size_t npk = picks.size();

    // one device block: picks | events | active list | states | residuals |
    // computed times | usable picks | their events | their scale flags
    const size_t b_pk = align256(sizeof(LocPick) * npk);
    const size_t b_ev = align256(sizeof(LocEvent) * (size_t)nev);
    const size_t b_ac = align256(sizeof(int) * (size_t)nev);
    const size_t b_st = align256(sizeof(LocState) * (size_t)nev);
    const size_t b_rs = align256(sizeof(float) * npk);
    const size_t b_sp = align256(sizeof(LocSPick) * npk);
    const size_t b_se = align256(sizeof(LocSEvent) * (size_t)nev);
    const size_t b_sc = align256(npk);
    const size_t bytes = b_pk + b_ev + b_ac + b_st + 2 * b_rs + b_sp + b_se + b_sc;
    char *blk = nullptr;
    int rc = dev_alloc(c, (void **)&blk, bytes);
    if (rc) return rc;
    size_t off = 0;
    auto carve = [&](size_t b) {
        char *p = blk + off;
        off += b;
        return p;
    };
    LocPick *d_pk = (LocPick *)carve(b_pk);
    LocEvent *d_ev = (LocEvent *)carve(b_ev);
    int *d_ac = (int *)carve(b_ac);
    LocState *d_st = (LocState *)carve(b_st);
    float *d_res = (float *)carve(b_rs);
    float *d_tps = (float *)carve(b_rs);
    LocSPick *d_sp = (LocSPick *)carve(b_sp);
    LocSEvent *d_se = (LocSEvent *)carve(b_se);
    unsigned char *d_sc = (unsigned char *)carve(b_sc);
    std::vector<LocSPick> spk;
    std::vector<LocSEvent> sev((size_t)nev);
    std::vector<unsigned char> ssc;
    spk.reserve(npk);
    ssc.reserve(npk);

    LocGrid g;
    memset(&g, 0, sizeof g);
    g.nx = P.nx; g.ny = P.ny; g.nz = P.nz;
    g.kmin = P.kmin; g.ndiv = P.ndiv; g.ndiv2 = P.ndiv2;
    g.vpvs = P.vpvs;
    g.tbytes = (unsigned)(4LL * P.nx * P.ny * P.nz);
    g.x0 = P.x0; g.y0 = P.y0; g.z0 = P.z0; g.df = P.df; g.dq = P.dq; g.h = P.h;

    std::vector<LocState> st((size_t)nev);
    std::vector<float> res(npk), tps(npk);
    std::vector<int> active;
    for (int e = 0; e < nev; e++) active.push_back(e);
    double kernel_ms = 0.0, node_ms = 0.0, node_bytes = 0.0;
    int npass = 0;
    bool ok = hipMemsetAsync(d_res, 0, sizeof(float) * npk, c->stream) == hipSuccess &&
              hipMemsetAsync(d_tps, 0, sizeof(float) * npk, c->stream) == hipSuccess;
    while (ok && !active.empty()) {
        // the pass's usable picks per event, and the variance terms no trial
        // changes, in the reference's float order (sphfdloc.c:781, 786-788)
        spk.clear();
        ssc.clear();
        for (int e : active) {
            LocEvent &E = events[(size_t)e];
            E.ngood = run[(size_t)e].ngood;
            LocSEvent &S = sev[(size_t)e];
            S.first = (int)spk.size();
            S.scaled = 0;
            float facs = 0.0f;
            for (int p = E.first; p < E.first + E.npick; p++) {
                const LocPick &k = picks[(size_t)p];
                if (!k.good) continue;
                spk.push_back(LocSPick{k.table, k.obs, k.pwt});
                ssc.push_back((unsigned char)k.scale);
                S.scaled |= k.scale;
                facs += k.pwt;
            }
            S.count = (int)spk.size() - S.first;
            S.facs = facs;
            const float avwt = facs / (float)E.ngood;
            S.den = (float)(E.ngood - 4) * avwt;
            S.pad = 0;
        }
        if (spk.empty()) {   // keep the device pointers valid for an all-flagged pass
            spk.push_back(LocSPick{picks[0].table, 0.0f, 0.0f});
            ssc.push_back(0);
        }
        ok = hipMemcpyAsync(d_pk, picks.data(), sizeof(LocPick) * npk, hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess &&
             hipMemcpyAsync(d_sp, spk.data(), sizeof(LocSPick) * spk.size(), hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess &&
             hipMemcpyAsync(d_se, sev.data(), sizeof(LocSEvent) * (size_t)nev, hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess &&
             hipMemcpyAsync(d_sc, ssc.data(), ssc.size(), hipMemcpyHostToDevice, c->stream) == hipSuccess &&
             hipMemcpyAsync(d_ev, events.data(), sizeof(LocEvent) * (size_t)nev, hipMemcpyHostToDevice,
                            c->stream) == hipSuccess &&
             hipMemcpyAsync(d_ac, active.data(), sizeof(int) * active.size(), hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess &&
             hipEventRecord(c->ev0, c->stream) == hipSuccess &&
             fd_launch_loc_pass(&g, d_pk, d_ev, d_sp, d_se, d_sc, d_ac, (int)active.size(), d_st, d_res, d_tps,
                                c->ev_a, c->ev_b, c->stream) == hipSuccess &&
             hipEventRecord(c->ev1, c->stream) == hipSuccess &&
             hipMemcpyAsync(st.data(), d_st, sizeof(LocState) * (size_t)nev, hipMemcpyDeviceToHost, c->stream) ==
                 hipSuccess &&
             hipMemcpyAsync(res.data(), d_res, sizeof(float) * npk, hipMemcpyDeviceToHost, c->stream) ==
                 hipSuccess &&
             hipMemcpyAsync(tps.data(), d_tps, sizeof(float) * npk, hipMemcpyDeviceToHost, c->stream) ==
                 hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
        if (!ok) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) kernel_ms += ms;
        if (hipEventElapsedTime(&ms, c->ev_a, c->ev_b) == hipSuccess) node_ms += ms;
        // the node search reads each table a usable pick of the pass names once
        // per searched node, however many events share it
        std::set<const float *> used;
        for (int e : active)
            for (int p = events[(size_t)e].first; p < events[(size_t)e].first + events[(size_t)e].npick; p++)
                if (picks[(size_t)p].good) used.insert(picks[(size_t)p].table);
        node_bytes += 4.0 * (double)used.size() * (double)P.nx * P.ny * std::max(0, P.nz - P.kmin);
        npass++;
        std::vector<int> next;
        for (int e : active) {
            EventRun &R = run[(size_t)e];
            const LocEvent &E = events[(size_t)e];
            const float avr = st[(size_t)e].avrmin;
            R.passes++;
            // drop outliers (sphfdloc.c:1080-1096)
            bool redo = false;
            int ngood = 0;
            for (int p = E.first; p < E.first + E.npick; p++)
                if (picks[(size_t)p].good) {
                    if (within(P, res[(size_t)p], avr, tps[(size_t)p])) {
                        ngood++;
                    } else {
                        picks[(size_t)p].good = 0;
                        redo = true;
                    }
                }
            R.ngood = ngood;
            if (ngood >= P.nthres) {
                if (!redo && !R.retried) {
                    // once, bring back picks the new hypocentre explains (1109-1126)
                    R.retried = true;
                    for (int p = E.first; p < E.first + E.npick; p++)
                        if (!picks[(size_t)p].good && within(P, res[(size_t)p], avr, tps[(size_t)p])) {
                            R.ngood++;
                            picks[(size_t)p].good = 1;
                            redo = true;
                        }
                }
                if (redo) {
                    next.push_back(e);
                    continue;
                }
            }
            R.done = true;
        }
        active.swap(next);
    }
    dev_free(c, blk, bytes);
    if (!ok) return FDGPU_EHIP;
    c->last_ms = kernel_ms;
    c->loc_timing[0] = npass;
    c->loc_timing[1] = kernel_ms;
    c->loc_timing[2] = node_ms;
    c->loc_timing[3] = node_bytes;
    for (int e = 0; e < nev; e++) {
        fdgpu_loc_event &E = ev[e];
        const LocEvent &L = events[(size_t)e];
        const LocState &S = st[(size_t)e];
        const EventRun &R = run[(size_t)e];
        for (int p = 0; p < E.npick; p++) {
            E.isgood[p] = picks[(size_t)(L.first + p)].good;
            E.res[p] = res[(size_t)(L.first + p)];
            E.tcalc[p] = tps[(size_t)(L.first + p)];
        }
        E.ngood = R.ngood;
        E.passes = R.passes;
        E.status = R.ngood < P.nthres ? 0 : S.stdmin <= P.stdmax ? 1 : 2;
        for (int q = 0; q < 3; q++) E.node[q] = S.node[q];
        E.stdmin = S.stdmin;
        E.avrmin = S.avrmin;
        E.exm = S.e[0];
        E.eym = S.e[1];
        E.ezm = S.e[2];
    }
    return FDGPU_OK;
}

int fdgpu_locate_timing(fdgpu_ctx *c, double *t, int n) {
    if (!c || !t || n < 0) return FDGPU_EINPUT;
    std::lock_guard<std::mutex> lk(c->mu);
    for (int q = 0; q < n && q < 4; q++) t[q] = c->loc_timing[q];
    return FDGPU_OK;
}
