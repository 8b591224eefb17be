// ray_kernel.hip -- gfx950 kernel of sphrayderv's ray tracing
// (src/sphrayderv.c:1306-2250) over resident travel-time tables.
//
// One lane per ray: a ray is a serial walk from cell to cell (each step
// depends on the last), so the parallelism is across rays.  The host orders
// a launch's rays by table, so the lanes of a wave walk one station's table
// and share its cache lines.  Each lane accumulates its ray's coarse-grid
// derivatives in its own row of a dense array (a segment's eight corners and
// the next segment's lie on few lines of that row, which stay in cache),
// exactly as the reference's dense per-ray
// array, in segment order.  Only the touched index range is scanned, and
// the scan leaves the column zero for the next launch.  The position's
// sines, cosines and arctangents are correctly rounded (cr_math.h).
//
// Compiled with -ffp-contract=off: every multiply and add rounds separately,
// as in the reference's x86-64 build.
#include <hip/hip_runtime.h>

#include "ray_kernels.h"

namespace {

constexpr int kThreads = 256;
// LDS for the grid's tables: up to this, else they stay in global memory
// (a 301 x 301 grid with a 203 x 203 x 105 coarse grid needs 21 KB)
constexpr size_t kLdsBytes = 48 * 1024;

// du[v] = (float)(du[v] + c) in this lane's row, and v's bit in the lane's
// row of touched-variable bits (word w holds variables 32w..32w+31).
struct RowAcc {
    float *du;
    unsigned *bits;
    int nvar, lo, hi;
    __device__ void add(int v, double c) {
        if (v < 0 || v >= nvar) return;
        float *p = du + v;
        const float old = *p;
        *p = (float)((double)old + c);
        // mark the variable on its first touch (a zero sum marks it again,
        // harmlessly); the OR is not waited for
        if (old == 0.0f) atomicOr(bits + (v >> 5), 1u << (v & 31));
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
};

__global__ __launch_bounds__(kThreads) void ray_kernel(fdray::RayGrid g, const fdray::RayIn *__restrict__ rays,
                                                       const float *const *__restrict__ tables, int n,
                                                       float *__restrict__ du, unsigned *__restrict__ bits, int nvar,
                                                       long long dstride, long long bstride,
                                                       RayResult *__restrict__ results,
                                                       unsigned long long *__restrict__ cursor,
                                                       int *__restrict__ idx, float *__restrict__ val,
                                                       long long capacity, int stage) {
    // the grid's tables (face sines and cosines, coarse coordinates, search
    // guesses) into LDS when the host sized them to fit (stage != 0): every
    // step of a walk reads them
    extern __shared__ double lds[];
    if (stage) {
        const int nt = 4 * (g.ny + 2) + 4 * (g.nx + 2), nc = g.nxc + g.nyc + g.nzc, nm = g.nx + g.ny + g.nz;
        double *t = lds;
        float *c = (float *)(t + nt);
        int *m = (int *)(c + nc);
        for (int q = threadIdx.x; q < nt; q += kThreads) t[q] = g.trig[q];
        for (int q = threadIdx.x; q < nc; q += kThreads)
            c[q] = q < g.nxc ? g.gx[q] : q < g.nxc + g.nyc ? g.gy[q - g.nxc] : g.gz[q - g.nxc - g.nyc];
        for (int q = threadIdx.x; q < nm; q += kThreads)
            m[q] = q < g.nx ? g.mx[q] : q < g.nx + g.ny ? g.my[q - g.nx] : g.mz[q - g.nx - g.ny];
        __syncthreads();
        g.trig = t;
        g.gx = c;
        g.gy = c + g.nxc;
        g.gz = c + g.nxc + g.nyc;
        g.mx = m;
        g.my = m + g.nx;
        g.mz = m + g.nx + g.ny;
    }
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= n) return;
    const fdray::RayIn in = rays[r];
    RowAcc acc{du + (long long)r * dstride, bits + (long long)r * bstride, nvar, nvar, -1};
    RayResult res;
    fdray::ray_trace<fdray::CrMath>(g, tables[in.table], in, res.o, acc);
    const bool ok = res.o.status == fdray::kRayOk;
    // count the nonzero touched variables, claim space, then write them in
    // index order and clear the column and its bits
    int k = 0;
    if (ok)
        for (int w = acc.lo >> 5; w <= (acc.hi >> 5); w++) {
            unsigned m = acc.bits[w];
            while (m) {
                const int v = (w << 5) + __builtin_ctz(m);
                m &= m - 1;
                k += acc.du[v] != 0.0f;
            }
        }
    long long off = -1;
    if (k > 0) {
        const unsigned long long o = atomicAdd(cursor, (unsigned long long)k);
        if ((long long)o + k <= capacity) off = (long long)o;
    }
    int wr = 0;
    for (int w = acc.lo >> 5; w <= (acc.hi >> 5); w++) {
        unsigned *bw = acc.bits + w;
        unsigned m = *bw;
        if (!m) continue;
        *bw = 0u;
        while (m) {
            const int v = (w << 5) + __builtin_ctz(m);
            m &= m - 1;
            float *p = acc.du + v;
            const float d = *p;
            *p = 0.0f;
            if (d != 0.0f && off >= 0) {
                idx[off + wr] = v;
                val[off + wr] = d;
                wr++;
            }
        }
    }
    res.o.ndu = ok ? k : 0;
    res.offset = k > 0 ? off : 0;
    results[r] = res;
}

}  // namespace

extern "C" hipError_t fd_launch_rays(const fdray::RayGrid *grid, const fdray::RayIn *rays, const float *const *tables,
                                     int n, float *du, unsigned *bits, int nvar, long long dstride, long long bstride,
                                     RayResult *results,
                                     unsigned long long *cursor, int *idx, float *val, long long capacity,
                                     hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const fdray::RayGrid &g = *grid;
    const size_t lds = sizeof(double) * (size_t)(4 * (g.ny + 2) + 4 * (g.nx + 2)) +
                       sizeof(float) * (size_t)(g.nxc + g.nyc + g.nzc) + sizeof(int) * (size_t)(g.nx + g.ny + g.nz);
    const int stage = lds <= kLdsBytes ? 1 : 0;
    hipLaunchKernelGGL(ray_kernel, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads),
                       stage ? lds : 0, stream, g, rays, tables, n, du, bits, nvar, dstride, bstride, results, cursor,
                       idx, val, capacity, stage);
    return hipGetLastError();
}

namespace {

constexpr int kPlaceThreads = 256;   // four rays per workgroup, one wave each

__global__ __launch_bounds__(kPlaceThreads) void ray_place_kernel(const RayResult *__restrict__ res,
                                                                  const long long *__restrict__ dst, int n,
                                                                  const int *__restrict__ idx,
                                                                  const float *__restrict__ val, int *out_idx,
                                                                  float *out_val) {
    const int k = blockIdx.x * (kPlaceThreads / 64) + (int)(threadIdx.x >> 6);
    if (k >= n) return;
    const int m = res[k].o.ndu;
    const long long from = res[k].offset, to = dst[k];
    if (m <= 0 || from < 0) return;
    for (int q = (int)(threadIdx.x & 63); q < m; q += 64) {
        out_idx[to + q] = idx[from + q];
        out_val[to + q] = val[from + q];
    }
}

}  // namespace

extern "C" hipError_t fd_launch_ray_place(const RayResult *results, const long long *dst, int n, const int *idx,
                                          const float *val, int *out_idx, float *out_val, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const int per = kPlaceThreads / 64;
    hipLaunchKernelGGL(ray_place_kernel, dim3((unsigned)((n + per - 1) / per)), dim3(kPlaceThreads), 0, stream,
                       results, dst, n, idx, val, out_idx, out_val);
    return hipGetLastError();
}
