/*
 * c2f_oracle.c -- CPU restatement of the coarse-to-fine model interpolation
 * of FDtomoC's c2f (src/c2f.c).  TEST INFRASTRUCTURE ONLY: the oracle that
 * the device path is checked against, and the CPU baseline of bench.py.
 *
 * Arithmetic follows the reference's types exactly: coarse node coordinates
 * and fine node coordinates are floats computed from doubles, the slowness
 * conversion rounds a double quotient to float, the interpolation is float
 * arithmetic in the reference's operation order, and the velocity written
 * is a float division.  Built without FP contraction (oracle/Makefile).
 */
#include "c2f_oracle.h"

#include <stdlib.h>

int oracle_c2f_dims(int nxc, int nyc, int nzc, const int *igridx, const int *igridy, const int *igridz, int *nx,
                    int *ny, int *nz) {
    if (nxc < 2 || nyc < 2 || nzc < 2) return -1;
    /* c2f.c:262-277: nx = 1 + sum of igridx */
    int a = 1, b = 1, c = 1;
    for (int i = 1; i < nxc; i++) a += igridx[i - 1];
    for (int i = 1; i < nyc; i++) b += igridy[i - 1];
    for (int i = 1; i < nzc; i++) c += igridz[i - 1];
    if (a < 1 || b < 1 || c < 1) return -1;
    *nx = a;
    *ny = b;
    *nz = c;
    return 0;
}

/* The coarse cell holding coordinate v on axis g (c2f.c:675-683): the first
 * node strictly beyond v, clamped to the last interval, minus one. */
static int cell(const float *g, int n, float v) {
    int i;
    for (i = 1; i < n; i++)
        if (g[i] > v) break;
    if (i == n) i--;
    return i - 1;
}

/* find_vel (c2f.c:673-727) over the slowness array sl (P then S). */
static float find_vel(const float *gx, const float *gy, const float *gz, int nxc, int nyc, int nzc, const float *sl,
                      float x, float y, float z, int iph) {
    const int nxyzc = nxc * nyc * nzc;
    int i = cell(gx, nxc, x);
    const float xi = gx[i];
    const float hx = gx[i + 1] - xi;
    const int j = cell(gy, nyc, y);
    const float yj = gy[j];
    const float hy = gy[j + 1] - yj;
    const int k = cell(gz, nzc, z);
    const float zk = gz[k];
    const float hz = gz[k + 1] - zk;
    const int nk = nxc * nyc * k;
    const int nj = nxc * j;
    const int nk2 = nxc * nyc * (k + 1);
    const int nj2 = nxc * (j + 1);
    const float fx = (x - xi) / hx;
    const float fy = (y - yj) / hy;
    const float fz = (z - zk) / hz;
    i = i + iph * nxyzc;
    float tv = 0.f;
    tv = (1.f - fx) * (1.f - fy) * (1.f - fz) * sl[nk + nj + i];
    tv = tv + fx * (1.f - fy) * (1.f - fz) * sl[nk + nj + i + 1];
    tv = tv + (1.f - fx) * fy * (1.f - fz) * sl[nk + nj2 + i];
    tv = tv + (1.f - fx) * (1.f - fy) * fz * sl[nk2 + nj + i];
    tv = tv + fx * fy * (1.f - fz) * sl[nk + nj2 + i + 1];
    tv = tv + fx * (1.f - fy) * fz * sl[nk2 + nj + i + 1];
    tv = tv + (1.f - fx) * fy * fz * sl[nk2 + nj2 + i];
    tv = tv + fx * fy * fz * sl[nk2 + nj2 + i + 1];
    return tv;
}

int oracle_c2f(int nxc, int nyc, int nzc, const int *igridx, const int *igridy, const int *igridz, double x0,
               double y0, double z0, double h, const float *vel, float *fine, int *nx, int *ny, int *nz) {
    int fx, fy, fz;
    if (oracle_c2f_dims(nxc, nyc, nzc, igridx, igridy, igridz, &fx, &fy, &fz)) return -1;
    *nx = fx;
    *ny = fy;
    *nz = fz;
    float *gx = malloc(sizeof(float) * (size_t)nxc);
    float *gy = malloc(sizeof(float) * (size_t)nyc);
    float *gz = malloc(sizeof(float) * (size_t)nzc);
    const long long nxyzc2 = 2LL * nxc * nyc * nzc;
    float *sl = malloc(sizeof(float) * (size_t)nxyzc2);
    if (!gx || !gy || !gz || !sl) {
        free(gx); free(gy); free(gz); free(sl);
        return -1;
    }
    /* c2f.c:261-277: node coordinates accumulated in float from the doubles */
    gx[0] = (float)x0;
    gy[0] = (float)y0;
    gz[0] = (float)z0;
    for (int i = 1; i < nxc; i++) gx[i] = (float)(gx[i - 1] + h * igridx[i - 1]);
    for (int i = 1; i < nyc; i++) gy[i] = (float)(gy[i - 1] + h * igridy[i - 1]);
    for (int i = 1; i < nzc; i++) gz[i] = (float)(gz[i - 1] + h * igridz[i - 1]);
    /* c2f.c:549-552: slowness, a double quotient rounded to float */
    for (long long i = 0; i < nxyzc2; i++) sl[i] = (float)(1. / (double)vel[i]);
    /* c2f.c:554-574: P then S, z, y, x; coordinates rounded to float */
    long long ij = 0;
    for (int n = 0; n <= 1; n++)
        for (int k = 0; k < fz; k++) {
            const float z = (float)(z0 + k * h);
            for (int j = 0; j < fy; j++) {
                const float y = (float)(y0 + j * h);
                for (int i = 0; i < fx; i++) {
                    const float x = (float)(x0 + i * h);
                    const float v = find_vel(gx, gy, gz, nxc, nyc, nzc, sl, x, y, z, n);
                    fine[ij++] = 1.f / v;
                }
            }
        }
    free(gx); free(gy); free(gz); free(sl);
    return 0;
}
