// coarse.cpp -- host side of the coarse-to-fine model API (include/fdtomoc_amd.h,
// fdgpu_coarse): the interpolation of FDtomoC's c2f (src/c2f.c) run on the
// device straight into resident P and S models.
//
// The host computes only what c2f computes before touching the model: the
// fine dimensions and the coarse node coordinates, accumulated in float from
// the spec's doubles (c2f.c:258-277).  The slowness conversion, the cell
// search, the trilinear interpolation and the velocity and slowness
// divisions run on the device (coarse_kernel.hip).
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "coarse_kernels.h"
#include "fdgpu_internal.h"

using namespace fdgpu_detail;

#define FD_HIP(call)                                  \
    do {                                              \
        hipError_t _e = (call);                       \
        if (_e != hipSuccess) return hip_fail(_e);    \
    } while (0)

namespace fdgpu_detail {

int coarse_check(const fdgpu_coarse *c, int &nx, int &ny, int &nz) {
    if (!c || !c->igridx || !c->igridy || !c->igridz || !c->velocity) return FDGPU_EINPUT;
    if (c->nxc < 2 || c->nyc < 2 || c->nzc < 2) return FDGPU_EINPUT;
    if (!(c->h > 0.0)) return FDGPU_EINPUT;
    long long a = 1, b = 1, d = 1;
    for (int i = 0; i < c->nxc - 1; i++) { if (c->igridx[i] < 1) return FDGPU_EINPUT; a += c->igridx[i]; }
    for (int i = 0; i < c->nyc - 1; i++) { if (c->igridy[i] < 1) return FDGPU_EINPUT; b += c->igridy[i]; }
    for (int i = 0; i < c->nzc - 1; i++) { if (c->igridz[i] < 1) return FDGPU_EINPUT; d += c->igridz[i]; }
    if (a > (1 << 24) || b > (1 << 24) || d > (1 << 24) || a * b * d >= (1LL << 31)) return FDGPU_EINPUT;
    nx = (int)a;
    ny = (int)b;
    nz = (int)d;
    return FDGPU_OK;
}

void coarse_coords(const fdgpu_coarse *co, std::vector<float> &g) {
    // c2f.c:258-277: node coordinates in float, accumulated from the doubles
    g.assign((size_t)(co->nxc + co->nyc + co->nzc), 0.f);
    float *gx = g.data(), *gy = gx + co->nxc, *gz = gy + co->nyc;
    gx[0] = (float)co->x0;
    gy[0] = (float)co->y0;
    gz[0] = (float)co->z0;
    for (int i = 1; i < co->nxc; i++) gx[i] = (float)(gx[i - 1] + co->h * co->igridx[i - 1]);
    for (int i = 1; i < co->nyc; i++) gy[i] = (float)(gy[i - 1] + co->h * co->igridy[i - 1]);
    for (int i = 1; i < co->nzc; i++) gz[i] = (float)(gz[i - 1] + co->h * co->igridz[i - 1]);
}

}  // namespace fdgpu_detail

namespace {

float bits_float(unsigned u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// Runs the kernels into the two models' slowness grids; sets their velocity
// ranges and copies the fine velocities to `fine` when it is not null.
int interpolate(fdgpu_ctx *c, const fdgpu_coarse *co, int nx, int ny, int nz, Model &mp, Model &ms, float *fine) {
    std::vector<float> g;
    coarse_coords(co, g);
    const long long ncoarse = 2LL * co->nxc * co->nyc * co->nzc;
    const long long nax = (long long)nx + ny + nz;
    const long long n = (long long)nx * ny * nz;
    // one scratch block: coarse values | coordinates | axis entries | range |
    // workgroup partial ranges | fine velocities
    const size_t b_co = align256(sizeof(float) * (size_t)ncoarse);
    const size_t b_g = align256(sizeof(float) * g.size());
    const size_t b_ax = align256(sizeof(C2fAxis) * (size_t)nax);
    const size_t b_rg = 256;
    const size_t b_part = align256(sizeof(unsigned) * 8 * (size_t)kC2fMaxBlocks);
    const size_t b_fine = fine ? align256(sizeof(float) * 2 * (size_t)n) : 0;
    const size_t bytes = b_co + b_g + b_ax + b_rg + b_part + b_fine;
    if (bytes > c->coarse_bytes) {
        FD_HIP(hipStreamSynchronize(c->stream));
        dev_free(c, c->d_coarse, c->coarse_bytes);
        c->d_coarse = nullptr;
        c->coarse_bytes = 0;
        const int rc = dev_alloc(c, (void **)&c->d_coarse, bytes);
        if (rc) return rc;
        c->coarse_bytes = bytes;
    }
    char *blk = c->d_coarse;
    float *d_co = (float *)blk;
    float *d_g = (float *)(blk + b_co);
    C2fAxis *d_ax = (C2fAxis *)(blk + b_co + b_g);
    unsigned *d_rg = (unsigned *)(blk + b_co + b_g + b_ax);
    unsigned *d_part = (unsigned *)(blk + b_co + b_g + b_ax + b_rg);
    float *d_fine = fine ? (float *)(blk + b_co + b_g + b_ax + b_rg + b_part) : nullptr;
    const unsigned rg0[8] = {0xffffffffu, 0u, 0u, 0u, 0xffffffffu, 0u, 0u, 0u};
    unsigned rg[8];
    bool ok = hipMemcpyAsync(d_co, co->velocity, sizeof(float) * (size_t)ncoarse, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(d_g, g.data(), sizeof(float) * g.size(), hipMemcpyHostToDevice, c->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(d_rg, rg0, sizeof rg0, hipMemcpyHostToDevice, c->stream) == hipSuccess;
    C2fGrid grid;
    grid.nxc = co->nxc; grid.nyc = co->nyc; grid.nzc = co->nzc;
    grid.gx = d_g; grid.gy = d_g + co->nxc; grid.gz = d_g + co->nxc + co->nyc;
    grid.nx = nx; grid.ny = ny; grid.nz = nz;
    grid.x0 = co->x0; grid.y0 = co->y0; grid.z0 = co->z0; grid.h = co->h;
    ok = ok && hipEventRecord(c->ev0, c->stream) == hipSuccess;
    ok = ok && fd_launch_c2f(&grid, d_co, d_ax, d_fine, mp.d_slow, ms.d_slow, d_part, d_rg, c->stream) == hipSuccess;
    ok = ok && hipEventRecord(c->ev1, c->stream) == hipSuccess;
    ok = ok && hipMemcpyAsync(rg, d_rg, sizeof rg, hipMemcpyDeviceToHost, c->stream) == hipSuccess;
    if (fine)
        ok = ok && hipMemcpyAsync(fine, d_fine, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, c->stream) ==
                       hipSuccess;
    ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    float ms_ = 0.f;
    if (ok && hipEventElapsedTime(&ms_, c->ev0, c->ev1) == hipSuccess) c->last_ms = ms_;
    if (!ok) {   // the slowness may be partly written: no skipping on either model
        mp.vmin = mp.vmax = 0.0;
        ms.vmin = ms.vmax = 0.0;
        return FDGPU_EHIP;
    }
    // velocity ranges bound the reverse-pass skip (fdprep::skip_safe); a
    // model with a non-positive or non-finite velocity never skips
    Model *m2[2] = {&mp, &ms};
    for (int p = 0; p < 2; p++) {
        const unsigned *r = rg + 4 * p;
        if (r[2] || r[0] > r[1]) {
            m2[p]->vmin = 0.0;
            m2[p]->vmax = 0.0;
        } else {
            m2[p]->vmin = bits_float(r[0]);
            m2[p]->vmax = bits_float(r[1]);
        }
    }
    return FDGPU_OK;
}

}  // namespace

int fdgpu_coarse_dims(const fdgpu_coarse *coarse, int *nx, int *ny, int *nz) {
    if (!nx || !ny || !nz) return FDGPU_EINPUT;
    return coarse_check(coarse, *nx, *ny, *nz);
}

int fdgpu_models_from_coarse(fdgpu_ctx *c, const fdgpu_coarse *coarse, int *model_p, int *model_s, float *fine) {
    if (!c || !model_p || !model_s) return FDGPU_EINPUT;
    int nx, ny, nz;
    int rc = coarse_check(coarse, nx, ny, nz);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;   // no allocation while the request loop runs
    FD_HIP(hipSetDevice(c->device));
    Model m[2];
    const size_t bytes = sizeof(float) * (size_t)nx * ny * nz;
    for (int p = 0; p < 2; p++) {
        m[p].nx = nx; m[p].ny = ny; m[p].nz = nz;
        m[p].bytes = bytes;
        rc = dev_alloc(c, (void **)&m[p].d_slow, bytes);
        if (rc) {
            if (p) dev_free(c, m[0].d_slow, bytes);
            return rc;
        }
    }
    rc = interpolate(c, coarse, nx, ny, nz, m[0], m[1], fine);
    if (rc) {
        dev_free(c, m[0].d_slow, bytes);
        dev_free(c, m[1].d_slow, bytes);
        return rc;
    }
    const int id_p = c->next_id++, id_s = c->next_id++;
    c->models[id_p] = m[0];
    c->models[id_s] = m[1];
    *model_p = id_p;
    *model_s = id_s;
    return FDGPU_OK;
}

int fdgpu_models_update_from_coarse(fdgpu_ctx *c, int model_p, int model_s, const fdgpu_coarse *coarse,
                                    float *fine) {
    if (!c || model_p == model_s) return FDGPU_EINPUT;
    int nx, ny, nz;
    int rc = coarse_check(coarse, nx, ny, nz);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->serve.on) return FDGPU_EBUSY;
    FD_HIP(hipSetDevice(c->device));
    auto ip = c->models.find(model_p), is = c->models.find(model_s);
    if (ip == c->models.end() || is == c->models.end()) return FDGPU_EINPUT;
    for (auto it : {ip, is})
        if (it->second.nx != nx || it->second.ny != ny || it->second.nz != nz) return FDGPU_EINPUT;
    // tables being solved read these grids: finish the stream's work first
    FD_HIP(hipStreamSynchronize(c->stream));
    return interpolate(c, coarse, nx, ny, nz, ip->second, is->second, fine);
}
