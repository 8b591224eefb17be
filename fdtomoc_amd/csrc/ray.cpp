// ray.cpp -- host side of sphrayderv's ray tracing over resident tables
// (include/fdtomoc_amd.h, fdgpu_rays_trace): the per-pick ray loop of
// src/sphrayderv.c:1069-2252 for a whole batch of picks in one launch per
// chunk of rays (ray_kernel.hip).
//
// The host orders the rays by table (lanes of a wave then walk the same
// table), computes the row and column sines and cosines of the grid's cell
// faces and each station's Cartesian point with its C library (the values
// the reference computes), and collects each ray's derivatives from the
// packed arrays the kernel appends to.  A ray whose derivatives did not fit
// the launch's packed arrays is traced again in a later launch (tracing is
// deterministic); every launch finishes at least one such ray, because the
// packed arrays hold any single ray's derivatives.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <unordered_map>
#include <numeric>
#include <tuple>
#include <vector>

#include "fdgpu_internal.h"
#include "ray_kernels.h"
#include "ray_prep.h"

using namespace fdgpu_detail;

namespace {

// Device memory for the lane columns of derivatives: at most this, or a
// quarter of what the budget has left (a lane's column is 4 bytes per
// coarse variable; every ray of a batch in one launch keeps the CUs busy).
constexpr size_t kColumnBytes = size_t(32) << 30;

bool grid_ok(const fdgpu_ray_grid &g) {
    if (g.nx < 2 || g.ny < 2 || g.nz < 2 || g.nxc < 2 || g.nyc < 2 || g.nzc < 2) return false;
    if (!g.gx || !g.gy || !g.gz) return false;
    if (!(g.df > 0.0) || !(g.dq > 0.0) || !(g.h > 0.0)) return false;
    if ((long long)g.nx * g.ny * g.nz >= (1LL << 31)) return false;
    if ((long long)g.nxc * g.nyc * g.nzc >= (1LL << 30)) return false;
    return g.max_segments >= 0;
}

}  // namespace

int fdgpu_rays_trace(fdgpu_ctx *c, const fdgpu_ray_grid *gp, const fdgpu_ray *rays, int n,
                     fdgpu_ray_result *results, int *du_index, float *du_value, long long du_cap) {
    if (!c || !gp || n < 0 || (n > 0 && (!rays || !results)) || du_cap < 0) return FDGPU_EINPUT;
    const fdgpu_ray_grid &G = *gp;
    if (!grid_ok(G)) return FDGPU_EINPUT;
    if (n == 0) return FDGPU_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;   // no allocation while the request loop runs
    if (hipSetDevice(c->device) != hipSuccess) return FDGPU_EHIP;

    // tables: each distinct one once, of the grid's dimensions
    std::unordered_map<int, int> slot;
    std::vector<const float *> tabs;
    std::vector<int> ray_slot((size_t)n);
    for (int r = 0; r < n; r++) {
        auto sl = slot.find(rays[r].table);
        if (sl == slot.end()) {
            auto it = c->tables.find(rays[r].table);
            if (it == c->tables.end() || !it->second.d_time) return FDGPU_EINPUT;
            const fdgpu_par &tp = it->second.par;
            if (tp.nx != G.nx || tp.ny != G.ny || tp.nz != G.nz) return FDGPU_EINPUT;
            sl = slot.emplace(rays[r].table, (int)tabs.size()).first;
            tabs.push_back(it->second.d_time);
        }
        ray_slot[(size_t)r] = sl->second;
    }
    // rays grouped by table (a wave's lanes then read one table), and within
    // a table by the event-station distance in cells, longest first (a wave
    // runs as long as its longest ray, so its lanes should walk alike): one
    // integer key per ray (table slot, then the distance's float bits
    // reversed), ties in input order
    std::vector<std::pair<unsigned long long, int>> keyed((size_t)n);
    for (int r = 0; r < n; r++) {
        const double di = (rays[r].ex - rays[r].xs) / G.df, dj = (rays[r].ey - rays[r].ys) / G.dq,
                     dk = (rays[r].ez - rays[r].zs) / G.h;
        const float reach = (float)(fabs(di) + fabs(dj) + fabs(dk));   // cells a walk crosses, about
        unsigned bits;
        memcpy(&bits, &reach, sizeof bits);   // non-negative: the bits order as the values
        keyed[(size_t)r] = {((unsigned long long)ray_slot[(size_t)r] << 32) | (0xffffffffu - bits), r};
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<int> order((size_t)n);
    for (int r = 0; r < n; r++) order[(size_t)r] = keyed[(size_t)r].second;
    // each station's Cartesian point once (a batch names few stations)
    struct PointKey {
        double x, y, z;
        bool operator==(const PointKey &o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct PointHash {
        size_t operator()(const PointKey &k) const {
            unsigned long long a, b, d;
            memcpy(&a, &k.x, 8);
            memcpy(&b, &k.y, 8);
            memcpy(&d, &k.z, 8);
            return (size_t)(a * 0x9e3779b97f4a7c15ull ^ (b + 0x632be59bd9b4e019ull) * 0xc2b2ae3d27d4eb4full ^ d);
        }
    };
    std::unordered_map<PointKey, std::tuple<float, float, float>, PointHash> points;
    std::vector<fdray::RayIn> in((size_t)n);
    for (int r = 0; r < n; r++) {
        const fdgpu_ray &R = rays[r];
        fdray::RayIn &q = in[(size_t)r];
        q.ex = R.ex; q.ey = R.ey; q.ez = R.ez;
        q.xs = R.xs; q.ys = R.ys; q.zs = R.zs;
        const PointKey key{R.xs, R.ys, R.zs};
        auto pt = points.find(key);
        if (pt == points.end()) {
            float x2, y2, z2;
            fdray::station_point(R.xs, R.ys, R.zs, x2, y2, z2);
            pt = points.emplace(key, std::make_tuple(x2, y2, z2)).first;
        }
        std::tie(q.x2, q.y2, q.z2) = pt->second;
        q.table = ray_slot[(size_t)r];
    }
    const std::vector<double> trig = fdray::face_trig(G.nx, G.ny, G.x0, G.y0, G.df, G.dq);
    std::vector<int> guess = fdray::coarse_guess(G.gx, G.nxc, G.x0, G.df, G.nx);
    const std::vector<int> gyv = fdray::coarse_guess(G.gy, G.nyc, G.y0, G.dq, G.ny),
                           gzv = fdray::coarse_guess(G.gz, G.nzc, G.z0, G.h, G.nz);
    guess.insert(guess.end(), gyv.begin(), gyv.end());
    guess.insert(guess.end(), gzv.begin(), gzv.end());
    const int nvar = G.ido1d ? G.nzc : G.nxc * G.nyc * G.nzc;

    // lanes per launch: the columns within their share of the budget
    // (the cached columns of an earlier call count as available)
    const size_t left = (c->budget > c->used ? (size_t)(c->budget - c->used) : 0) + c->ray_cols_bytes;
    const size_t col_budget = std::min(kColumnBytes, left / 4);
    long long lanes = (long long)(col_budget / (sizeof(float) * (size_t)nvar + sizeof(unsigned) * (size_t)((nvar + 31) / 32)));
    lanes = std::min<long long>(lanes, n);
    if (lanes < 1) return FDGPU_ENOMEM;
    if (lanes > 256) lanes = lanes / 256 * 256;
    // each lane's row of derivatives and of touched bits, 128-byte aligned
    const long long dstride = ((long long)nvar + 31) / 32 * 32;
    const long long bstride = (((long long)nvar + 31) / 32 + 31) / 32 * 32;
    // packed derivatives of a launch: about 512 per ray, never fewer than one ray's worst case
    const long long capacity = std::max<long long>(nvar, std::min<long long>(lanes * 512, 1LL << 26));

    const size_t b_g = align256(sizeof(float) * (size_t)(G.nxc + G.nyc + G.nzc));
    const size_t b_t = align256(sizeof(double) * trig.size());
    const size_t b_m = align256(sizeof(int) * guess.size());
    const size_t b_p = align256(sizeof(const float *) * tabs.size());
    const size_t b_r = align256(sizeof(fdray::RayIn) * (size_t)lanes);
    const size_t b_o = align256(sizeof(RayResult) * (size_t)lanes);
    const size_t b_d = align256(sizeof(float) * (size_t)dstride * (size_t)lanes);
    const size_t b_b = align256(sizeof(unsigned) * (size_t)bstride * (size_t)lanes);
    const size_t b_c = 256;
    const size_t b_i = align256(sizeof(int) * (size_t)capacity);
    const size_t b_v = align256(sizeof(float) * (size_t)capacity);
    // the lane columns: cached in the context between calls (every launch
    // leaves them zero), grown when a call needs more
    bool ok = true;
    if (c->ray_cols_bytes < b_d + b_b) {
        release_ray_columns(c);
        int rc = dev_alloc(c, (void **)&c->ray_cols, b_d + b_b);
        if (rc) return rc;
        c->ray_cols_bytes = b_d + b_b;
        ok = hipMemsetAsync(c->ray_cols, 0, c->ray_cols_bytes, c->stream) == hipSuccess;
    }
    const size_t blk_bytes = b_g + b_t + b_m + b_p + b_r + b_o + b_c + b_i + b_v;
    if (c->ray_blk_bytes < blk_bytes) {
        (void)hipStreamSynchronize(c->stream);
        dev_free(c, c->ray_blk, c->ray_blk_bytes);
        c->ray_blk = nullptr;
        c->ray_blk_bytes = 0;
        int rc = dev_alloc(c, (void **)&c->ray_blk, blk_bytes);
        if (rc) return rc;
        c->ray_blk_bytes = blk_bytes;
    }
    char *blk = c->ray_blk;
    size_t off = 0;
    auto carve = [&](size_t b) {
        char *p = blk + off;
        off += b;
        return p;
    };
    float *d_g = (float *)carve(b_g);
    double *d_trig = (double *)carve(b_t);
    int *d_guess = (int *)carve(b_m);
    const float **d_tabs = (const float **)carve(b_p);
    fdray::RayIn *d_in = (fdray::RayIn *)carve(b_r);
    RayResult *d_res = (RayResult *)carve(b_o);
    float *d_du = (float *)c->ray_cols;
    unsigned *d_bits = (unsigned *)(c->ray_cols + b_d);
    unsigned long long *d_cursor = (unsigned long long *)carve(b_c);
    int *d_idx = (int *)carve(b_i);
    float *d_val = (float *)carve(b_v);

    std::vector<float> gxyz;
    gxyz.insert(gxyz.end(), G.gx, G.gx + G.nxc);
    gxyz.insert(gxyz.end(), G.gy, G.gy + G.nyc);
    gxyz.insert(gxyz.end(), G.gz, G.gz + G.nzc);
    fdray::RayGrid dg{G.nx, G.ny, G.nz, G.nxc, G.nyc, G.nzc, G.ido1d ? 1 : 0,
                      G.max_segments > 0 ? G.max_segments : 8 * (G.nx + G.ny + G.nz),
                      G.x0, G.y0, G.z0, G.df, G.dq, G.h, d_g, d_g + G.nxc, d_g + G.nxc + G.nyc, d_trig,
                      d_guess, d_guess + G.nx, d_guess + G.nx + G.ny};
    ok = ok &&
         hipMemcpyAsync(d_g, gxyz.data(), sizeof(float) * gxyz.size(), hipMemcpyHostToDevice, c->stream) ==
                  hipSuccess &&
              hipMemcpyAsync(d_trig, trig.data(), sizeof(double) * trig.size(), hipMemcpyHostToDevice, c->stream) ==
                  hipSuccess &&
              hipMemcpyAsync(d_guess, guess.data(), sizeof(int) * guess.size(), hipMemcpyHostToDevice, c->stream) ==
                  hipSuccess &&
              hipMemcpyAsync(d_tabs, tabs.data(), sizeof(const float *) * tabs.size(), hipMemcpyHostToDevice,
                             c->stream) == hipSuccess;

    // each ray's result and derivatives, as the launches deliver them
    std::vector<fdray::RayOut> outs((size_t)n);
    std::vector<long long> start((size_t)n, 0);
    std::vector<int> h_idx;
    std::vector<float> h_val;
    std::vector<fdray::RayIn> chunk;
    std::vector<int> chunk_ray;
    std::vector<RayResult> got;
    std::vector<int> pending = order, retry;
    double ms_total = 0.0;
    bool one_launch = false;
    while (ok && !pending.empty()) {
        retry.clear();
        for (size_t p0 = 0; ok && p0 < pending.size(); p0 += (size_t)lanes) {
            const int m = (int)std::min<size_t>((size_t)lanes, pending.size() - p0);
            chunk.assign((size_t)m, fdray::RayIn{});
            chunk_ray.assign(pending.begin() + (long)p0, pending.begin() + (long)p0 + m);
            for (int k = 0; k < m; k++) chunk[(size_t)k] = in[(size_t)chunk_ray[(size_t)k]];
            ok = hipMemcpyAsync(d_in, chunk.data(), sizeof(fdray::RayIn) * (size_t)m, hipMemcpyHostToDevice,
                                c->stream) == hipSuccess &&
                 hipMemsetAsync(d_cursor, 0, sizeof(unsigned long long), c->stream) == hipSuccess &&
                 hipEventRecord(c->ev0, c->stream) == hipSuccess &&
                 fd_launch_rays(&dg, d_in, d_tabs, m, d_du, d_bits, nvar, dstride, bstride, d_res, d_cursor, d_idx,
                                d_val, capacity, c->stream) == hipSuccess &&
                 hipEventRecord(c->ev1, c->stream) == hipSuccess;
            got.resize((size_t)m);
            unsigned long long used = 0;
            ok = ok &&
                 hipMemcpyAsync(got.data(), d_res, sizeof(RayResult) * (size_t)m, hipMemcpyDeviceToHost, c->stream) ==
                     hipSuccess &&
                 hipMemcpyAsync(&used, d_cursor, sizeof(used), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
                 hipStreamSynchronize(c->stream) == hipSuccess;
            if (!ok) break;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) ms_total += ms;
            // every ray in this one launch (only the first pass holds all n),
            // every derivative fitted: they are placed in output order on
            // the device and copied once (below)
            bool fits = m == n;
            for (int k = 0; fits && k < m; k++)
                if (got[(size_t)k].o.ndu > 0 && got[(size_t)k].offset < 0) fits = false;
            if (fits) {
                one_launch = true;
                for (int k = 0; k < m; k++) outs[(size_t)chunk_ray[(size_t)k]] = got[(size_t)k].o;
                continue;
            }
            const long long filled = std::min<long long>((long long)used, capacity);
            const size_t base = h_idx.size();
            h_idx.resize(base + (size_t)filled);
            h_val.resize(base + (size_t)filled);
            if (filled > 0)
                ok = hipMemcpy(h_idx.data() + base, d_idx, sizeof(int) * (size_t)filled, hipMemcpyDeviceToHost) ==
                         hipSuccess &&
                     hipMemcpy(h_val.data() + base, d_val, sizeof(float) * (size_t)filled, hipMemcpyDeviceToHost) ==
                         hipSuccess;
            for (int k = 0; k < m; k++) {
                const RayResult &R = got[(size_t)k];
                const int r = chunk_ray[(size_t)k];
                if (R.o.ndu > 0 && R.offset < 0) {
                    retry.push_back(r);   // its derivatives did not fit: trace it again
                    continue;
                }
                outs[(size_t)r] = R.o;
                start[(size_t)r] = (long long)base + (R.o.ndu > 0 ? R.offset : 0);
            }
        }
        pending.swap(retry);
    }
    if (!ok) {
        release_ray_columns(c);   // a failed launch may have left them dirty
        return FDGPU_EHIP;
    }
    c->last_ms = ms_total;

    // results in input order; the derivatives packed in the same order
    long long total = 0;
    for (int r = 0; r < n; r++) {
        const fdray::RayOut &o = outs[(size_t)r];
        fdgpu_ray_result &q = results[r];
        q.status = o.status;
        q.nseg = o.nseg;
        q.ndu = o.ndu;
        q.first_in_cube = o.first_in_cube;
        q.tev = o.tev;
        q.dtdx = o.dtdx; q.dtdy = o.dtdy; q.dtdz = o.dtdz;
        fdray::take_off(o, q.ai, q.az);
        q.rdev = o.rdev;
        q.offset = total;
        total += o.ndu;
    }
    if (total > du_cap || (total > 0 && (!du_index || !du_value))) return FDGPU_ESPACE;
    if (one_launch) {
        // the launch's packed derivatives to their output offsets on the
        // device, then one copy into the caller's arrays
        if (total == 0) return FDGPU_OK;
        const size_t b_dst = align256(sizeof(long long) * (size_t)n);
        const size_t need = b_dst + align256(sizeof(int) * (size_t)total) + align256(sizeof(float) * (size_t)total);
        if (c->ray_out_bytes < need) {
            (void)hipStreamSynchronize(c->stream);
            dev_free(c, c->ray_out, c->ray_out_bytes);
            c->ray_out = nullptr;
            c->ray_out_bytes = 0;
            int rc = dev_alloc(c, (void **)&c->ray_out, need);
            if (rc) return rc;
            c->ray_out_bytes = need;
        }
        long long *d_dst = (long long *)c->ray_out;
        int *d_oidx = (int *)(c->ray_out + b_dst);
        float *d_oval = (float *)(c->ray_out + b_dst + align256(sizeof(int) * (size_t)total));
        std::vector<long long> dst((size_t)n);
        for (int k = 0; k < n; k++) dst[(size_t)k] = results[chunk_ray[(size_t)k]].offset;
        ok = hipMemcpyAsync(d_dst, dst.data(), sizeof(long long) * (size_t)n, hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess &&
             fd_launch_ray_place(d_res, d_dst, n, d_idx, d_val, d_oidx, d_oval, c->stream) == hipSuccess &&
             hipMemcpyAsync(du_index, d_oidx, sizeof(int) * (size_t)total, hipMemcpyDeviceToHost, c->stream) ==
                 hipSuccess &&
             hipMemcpyAsync(du_value, d_oval, sizeof(float) * (size_t)total, hipMemcpyDeviceToHost, c->stream) ==
                 hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
        return ok ? FDGPU_OK : FDGPU_EHIP;
    }
    for (int r = 0; r < n; r++) {
        const int k = outs[(size_t)r].ndu;
        if (k <= 0) continue;
        memcpy(du_index + results[r].offset, h_idx.data() + start[(size_t)r], sizeof(int) * (size_t)k);
        memcpy(du_value + results[r].offset, h_val.data() + start[(size_t)r], sizeof(float) * (size_t)k);
    }
    return FDGPU_OK;
}
