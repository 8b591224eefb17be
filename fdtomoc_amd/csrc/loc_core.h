// loc_core.h -- device functions of sphfdloc's grid-search location
// (src/sphfdloc.c:752-1052, find_time 1305-1345): the trilinear table time,
// a trial point's weighted residual statistics, and the two refinement
// boxes.  Double table times and weights, float residuals and sums, in the
// reference's operation order; compiled without FP contraction.
#pragma once

#include <hip/hip_runtime.h>

#include "loc_kernels.h"

namespace locdev {

constexpr unsigned long long kNone = ~0ull;

// find_time's cell (sphfdloc.c:1306-1335): the clamped cell holding the
// point, as the offset of its low corner, and the eight corner weights in the
// reference's term order, each the product of its three factors left to
// right.  The weights depend only on the point, so a trial computes them
// once for all of its picks.
struct Cell {
    long long base;
    double w[8];
};

__device__ __forceinline__ Cell cell_of(const LocGrid &g, double x, double y, double z) {
    int i = (int)((x - g.x0) / g.df);
    int j = (int)((y - g.y0) / g.dq);
    int k = (int)((z - g.z0) / g.h);
    i = max(0, min(i, g.nx - 2));
    j = max(0, min(j, g.ny - 2));
    k = max(0, min(k, g.nz - 2));
    const double fx = (x - (g.x0 + g.df * i)) / g.df;
    const double fy = (y - (g.y0 + g.dq * j)) / g.dq;
    const double fz = (z - (g.z0 + g.h * k)) / g.h;
    const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
    Cell c;
    c.base = (long long)g.nx * g.ny * k + (long long)g.nx * j + i;
    c.w[0] = gx * gy * gz;
    c.w[1] = fx * gy * gz;
    c.w[2] = gx * fy * gz;
    c.w[3] = gx * gy * fz;
    c.w[4] = fx * fy * gz;
    c.w[5] = fx * gy * fz;
    c.w[6] = gx * fy * fz;
    c.w[7] = fx * fy * fz;
    return c;
}

// The table time at the cell's point: the eight terms summed in order
// (sphfdloc.c:1337-1344).
__device__ __forceinline__ double interp(const float *__restrict__ t, const Cell &c, long long sy, long long sz) {
    const __attribute__((address_space(1))) float *p = (const __attribute__((address_space(1))) float *)t + c.base;
    double tp = c.w[0] * (double)p[0];
    tp = tp + c.w[1] * (double)p[1];
    tp = tp + c.w[2] * (double)p[sy];
    tp = tp + c.w[3] * (double)p[sz];
    tp = tp + c.w[4] * (double)p[sy + 1];
    tp = tp + c.w[5] * (double)p[sz + 1];
    tp = tp + c.w[6] * (double)p[sz + sy];
    tp = tp + c.w[7] * (double)p[sz + sy + 1];
    return tp;
}

// A pick's residual at a trial point from its table time (sphfdloc.c:768-772).
__device__ __forceinline__ float residual(const LocGrid &g, const LocPick &p, double t) {
    if (p.scale) t *= (double)g.vpvs;
    return (float)((double)p.obs - t);
}

struct Sums {
    float avres, rsq, facs;
};

// A usable pick's terms (sphfdloc.c:773-781).
__device__ __forceinline__ void add(Sums &s, float r, float w) {
    const float r1 = r * w;
    const float r2 = r1 * r;
    s.avres += r1;
    s.rsq += r2;
    s.facs += w;
}

// The trial's deviation (sphfdloc.c:786-790).
__device__ __forceinline__ float deviation(const Sums &s, int ngood) {
    const float avwt = s.facs / (float)ngood;
    const float var = (s.rsq - s.avres * s.avres / s.facs) / ((float)(ngood - 4) * avwt);
    return sqrtf(fabsf(var));
}

// The same from the event's fixed terms: facs and the divisor (ngood - 4) *
// avwt do not depend on the trial point.
__device__ __forceinline__ float deviation(float avres, float rsq, float facs, float den) {
    const float var = (rsq - avres * avres / facs) / den;
    return sqrtf(fabsf(var));
}

// A usable pick's two sums from its residual.
__device__ __forceinline__ void add2(float &avres, float &rsq, float r, float w) {
    const float r1 = r * w;
    const float r2 = r1 * r;
    avres += r1;
    rsq += r2;
}

// Residual of an unscaled pick: (float)((double)obs - (double)t) of the
// reference is one float subtraction, since rounding the exact difference of
// two floats first to double (53 >= 2 * 24 + 2 bits) and then to float gives
// the single float rounding (Figueroa's double-rounding bound).
__device__ __forceinline__ float residual_unscaled(float obs, float t) { return obs - t; }

// A pick's table as a buffer resource: its base in scalar registers and its
// size as the bound (raw 32-bit access; dword 3 as for every gfx9 target).
// A load then takes only the lane's byte offset: a refinement point's eight
// corner loads per pick need no vector address arithmetic.  The host keeps tables below 2^30 nodes, so
// byte offsets and sizes fit 32 bits.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const float *t, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)t, (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, unsigned off4) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off4, 0, 0));
}

// Table value through the global address space (a pick's table pointer is
// read from memory, so the compiler would otherwise use a flat load).  The
// node search takes one value per table, so building a buffer resource per
// pick would cost more scalar instructions than the address add it saves.
__device__ __forceinline__ float gload(const float *t, long long n) {
    return *((const __attribute__((address_space(1))) float *)t + n);
}

// The values of nodes n and n + 1 (n even: 8-byte aligned in a table) in one load.
typedef float float_pair __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float_pair gload2(const float *t, long long n) {
    return *(const __attribute__((address_space(1))) float_pair *)((const __attribute__((address_space(1))) float *)t +
                                                                   n);
}

// The byte offsets of a cell's eight corners in find_time's term order,
// once per point for all of its picks' tables.
__device__ __forceinline__ void corner_offsets(const Cell &c, long long sy, long long sz, unsigned o[8]) {
    const unsigned b = (unsigned)c.base * 4u, y = (unsigned)sy * 4u, z = (unsigned)sz * 4u;
    o[0] = b;
    o[1] = b + 4u;
    o[2] = b + y;
    o[3] = b + z;
    o[4] = b + y + 4u;
    o[5] = b + z + 4u;
    o[6] = b + z + y;
    o[7] = b + z + y + 4u;
}

// interp() through a table's buffer resource: the same terms in the same order.
__device__ __forceinline__ double interp_b(__amdgpu_buffer_rsrc_t r, const Cell &c, const unsigned o[8]) {
    double tp = c.w[0] * (double)bload(r, o[0]);
    tp = tp + c.w[1] * (double)bload(r, o[1]);
    tp = tp + c.w[2] * (double)bload(r, o[2]);
    tp = tp + c.w[3] * (double)bload(r, o[3]);
    tp = tp + c.w[4] * (double)bload(r, o[4]);
    tp = tp + c.w[5] * (double)bload(r, o[5]);
    tp = tp + c.w[6] * (double)bload(r, o[6]);
    tp = tp + c.w[7] * (double)bload(r, o[7]);
    return tp;
}

// Ordering key of a trial that beats `bound` (the strict comparison of
// sphfdloc.c:793): deviations are +0 or positive, so their bit patterns
// order as their values, and among equal deviations the lower scan index
// (the one the reference's loops reach first) wins.
__device__ __forceinline__ unsigned long long key_of(float dev, float bound, unsigned idx) {
    return dev < bound ? ((unsigned long long)__float_as_uint(dev) << 32) | idx : kNone;
}

// A refinement box: n[a] points per axis from o[a] in steps d[a].
struct Box {
    double o[3], d[3];
    int n[3];
};

// The first refinement's box around the node (sphfdloc.c:844-884).
__device__ __forceinline__ Box box_first(const LocGrid &g, const int node[3]) {
    const int n1 = g.ndiv + 1, n2 = n1 + g.ndiv;
    const double org[3] = {g.x0, g.y0, g.z0}, sp[3] = {g.df, g.dq, g.h};
    const int nn[3] = {g.nx, g.ny, g.nz};
    Box b;
    for (int a = 0; a < 3; a++) {
        const int lo = a == 2 ? g.kmin : 0;
        int s = node[a] - 1;
        b.n[a] = n2;
        if (s < lo) {
            s = node[a];
            b.n[a] = n1;
        }
        if (node[a] == nn[a] - 1) b.n[a] = n1 - 1;
        b.o[a] = org[a] + sp[a] * s;
        b.d[a] = sp[a] / g.ndiv;
    }
    return b;
}

// The second refinement's box around the hypocentre (sphfdloc.c:958-1001);
// the far faces are the grid's far corner rounded to float (502-504).
__device__ __forceinline__ Box box_second(const LocGrid &g, const double e[3]) {
    const int n1 = g.ndiv2 + 1, n2 = n1 + g.ndiv2;
    const double sp[3] = {g.df, g.dq, g.h};
    const float far[3] = {(float)(g.x0 + (g.nx - 1) * g.df), (float)(g.y0 + (g.ny - 1) * g.dq),
                          (float)(g.z0 + (g.nz - 1) * g.h)};
    const double low[3] = {g.x0, g.y0, g.z0 + g.kmin * g.h};
    Box b;
    for (int a = 0; a < 3; a++) {
        const double d1 = sp[a] / g.ndiv;
        b.o[a] = e[a] - d1;
        b.n[a] = n2;
        if (b.o[a] <= low[a]) {
            b.o[a] = e[a];
            b.n[a] = n1;
        }
        if (e[a] + d1 > (double)far[a]) b.n[a] = n1 - 1;
        b.d[a] = d1 / g.ndiv2;
    }
    return b;
}

__device__ __forceinline__ unsigned box_points(const Box &b) {
    if (b.n[0] <= 0 || b.n[1] <= 0 || b.n[2] <= 0) return 0u;
    return (unsigned)b.n[0] * (unsigned)b.n[1] * (unsigned)b.n[2];
}

// Point q of a box in the reference's loop order (z outer, x inner); each
// coordinate is the box origin plus step times count (sphfdloc.c:887-892).
__device__ __forceinline__ void box_point(const Box &b, unsigned q, double p[3]) {
    const unsigned i = q % (unsigned)b.n[0];
    const unsigned r = q / (unsigned)b.n[0];
    const unsigned j = r % (unsigned)b.n[1];
    const unsigned k = r / (unsigned)b.n[1];
    p[0] = b.o[0] + b.d[0] * (int)i;
    p[1] = b.o[1] + b.d[1] * (int)j;
    p[2] = b.o[2] + b.d[2] * (int)k;
}

}  // namespace locdev
