/*
 * loc_oracle.c -- CPU restatement of sphfdloc's grid-search location
 * (src/sphfdloc.c:752-1130, find_time 1305-1345).  Test checker only; the
 * header states how it is pinned.
 *
 * Arithmetic follows the reference's types and order: table times and the
 * trilinear weights in double, residuals stored as float, the weighted sums,
 * the variance and the deviation in float; no contraction (built with
 * -ffp-contract=off).
 */
#include "loc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

double oracle_loc_find_time(const oracle_loc_grid *g, const float *t, double x, double y, double z) {
    /* cell holding the point, clamped to the grid's last cell (sphfdloc.c:1306-1324) */
    int i = (int)((x - g->x0) / g->df);
    int j = (int)((y - g->y0) / g->dq);
    int k = (int)((z - g->z0) / g->h);
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    if (k < 0) k = 0;
    if (i > g->nx - 2) i = g->nx - 2;
    if (j > g->ny - 2) j = g->ny - 2;
    if (k > g->nz - 2) k = g->nz - 2;
    const double fx = (x - (g->x0 + g->df * i)) / g->df;
    const double fy = (y - (g->y0 + g->dq * j)) / g->dq;
    const double fz = (z - (g->z0 + g->h * k)) / g->h;
    const long sx = 1, sy = g->nx, sz = (long)g->nx * g->ny;
    const float *c = t + sz * k + sy * j + i;
    /* the eight corners in the reference's order (sphfdloc.c:1337-1344);
     * (1.f - f) is a double since f is */
    const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
    double tp = gx * gy * gz * c[0];
    tp = tp + fx * gy * gz * c[sx];
    tp = tp + gx * fy * gz * c[sy];
    tp = tp + gx * gy * fz * c[sz];
    tp = tp + fx * fy * gz * c[sy + sx];
    tp = tp + fx * gy * fz * c[sz + sx];
    tp = tp + gx * fy * fz * c[sz + sy];
    tp = tp + fx * fy * fz * c[sz + sy + sx];
    return tp;
}

/* One event's picks and the running minimum of one pass. */
typedef struct {
    const oracle_loc_grid *g;
    const float *const *tab;
    int npick;
    const float *obs, *pwt;
    const char *ph;
    const int *good;
    int ngood;
    double *tp;             /* per pick: table time at the trial point */
    float *res;             /* per pick: residual at the trial point */
    float *resmin, *tps;
    float stdmin, avrmin;
} pass_t;

/* Deviation of the trial point whose table times are in P->tp
 * (sphfdloc.c:766-792); keeps it when it beats the minimum so far (793-803). */
static int trial(pass_t *P) {
    float avres = 0.0f, rsq = 0.0f, facs = 0.0f;
    for (int p = 0; p < P->npick; p++) {
        double t = P->tp[p];
        if (P->g->ivs == 0 && P->ph[p] == 'S') t *= P->g->vpvs;
        P->res[p] = (float)(P->obs[p] - t);
        if (P->good[p]) {
            const float w = P->pwt[p];
            const float r1 = P->res[p] * w;
            const float r2 = r1 * P->res[p];
            avres += r1;
            rsq += r2;
            facs += w;
        }
    }
    const float avwt = facs / (float)P->ngood;
    const float var = (rsq - avres * avres / facs) / ((float)(P->ngood - 4) * avwt);
    const float dev = sqrtf(fabsf(var));
    if (!(dev < P->stdmin)) return 0;
    P->stdmin = dev;
    P->avrmin = avres / facs;
    for (int p = 0; p < P->npick; p++) {
        P->resmin[p] = P->res[p];
        P->tps[p] = P->obs[p] - P->res[p];
    }
    return 1;
}

/* One refinement box: n[a] points per axis from o[a] in steps d[a], z outer
 * and x inner (sphfdloc.c:886-940, 1002-1052); moves the hypocentre to the
 * best point when one beats the minimum. */
static void refine(pass_t *P, const double o[3], const double d[3], const int n[3], double e[3]) {
    for (int k = 0; k < n[2]; k++) {
        const double zp = o[2] + d[2] * k;
        for (int j = 0; j < n[1]; j++) {
            const double yp = o[1] + d[1] * j;
            for (int i = 0; i < n[0]; i++) {
                const double xp = o[0] + d[0] * i;
                for (int p = 0; p < P->npick; p++) P->tp[p] = oracle_loc_find_time(P->g, P->tab[p], xp, yp, zp);
                if (trial(P)) {
                    e[0] = xp;
                    e[1] = yp;
                    e[2] = zp;
                }
            }
        }
    }
}

/* One search pass (sphfdloc.c:752-1052): node search from layer kmin, then
 * the two refinements. */
static void search(pass_t *P, int node[3], double e[3]) {
    const oracle_loc_grid *g = P->g;
    const long nxy = (long)g->nx * g->ny;
    int b[3] = {0, 0, 0};
    for (int k = g->kmin; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
                const long n = nxy * k + (long)g->nx * j + i;
                for (int p = 0; p < P->npick; p++) P->tp[p] = P->tab[p][n];
                if (trial(P)) {
                    b[0] = i;
                    b[1] = j;
                    b[2] = k;
                }
            }
    memcpy(node, b, sizeof b);
    const double org[3] = {g->x0, g->y0, g->z0};
    const double sp[3] = {g->df, g->dq, g->h};
    const int nn[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; a++) e[a] = org[a] + b[a] * sp[a];

    /* first refinement: one node either side of the best node in steps of
     * 1/ndiv, one-sided at the grid's faces (sphfdloc.c:844-884); the low
     * face of depth is layer kmin, and at a high face the box stops a step
     * short of the face node */
    {
        const int n1 = g->ndiv + 1, n2 = n1 + g->ndiv;
        const double d[3] = {g->df / g->ndiv, g->dq / g->ndiv, g->h / g->ndiv};
        double o[3];
        int n[3];
        for (int a = 0; a < 3; a++) {
            const int lo = a == 2 ? g->kmin : 0;
            int s = b[a] - 1;
            n[a] = n2;
            if (s < lo) {
                s = b[a];
                n[a] = n1;
            }
            if (b[a] == nn[a] - 1) n[a] = n1 - 1;
            o[a] = org[a] + sp[a] * s;
        }
        refine(P, o, d, n, e);
    }
    /* second refinement: one first-refinement step either side of the
     * hypocentre in steps of 1/ndiv2 of it (sphfdloc.c:958-1001); the faces
     * are tested in double against the origin (depth: layer kmin) and
     * against the far corner rounded to float (502-504) */
    {
        const int n1 = g->ndiv2 + 1, n2 = n1 + g->ndiv2;
        const double d1[3] = {g->df / g->ndiv, g->dq / g->ndiv, g->h / g->ndiv};
        const double d[3] = {d1[0] / g->ndiv2, d1[1] / g->ndiv2, d1[2] / g->ndiv2};
        const float far[3] = {(float)(g->x0 + (g->nx - 1) * g->df), (float)(g->y0 + (g->ny - 1) * g->dq),
                              (float)(g->z0 + (g->nz - 1) * g->h)};
        const double low[3] = {g->x0, g->y0, g->z0 + g->kmin * g->h};
        double o[3];
        int n[3];
        for (int a = 0; a < 3; a++) {
            o[a] = e[a] - d1[a];
            n[a] = n2;
            if (o[a] <= low[a]) {
                o[a] = e[a];
                n[a] = n1;
            }
            if (e[a] + d1[a] > far[a]) n[a] = n1 - 1;
        }
        refine(P, o, d, n, e);
    }
}

/* Outlier test of one pick at the hypocentre (sphfdloc.c:1085-1087, 1115-1117). */
static int within(const oracle_loc_grid *g, float resmin, float avrmin, float tps) {
    const float absres = fabsf(resmin - avrmin);
    const float percres = 100.0f * absres / tps;
    return absres <= g->resthres || percres <= g->resthrep;
}

int oracle_locate(const oracle_loc_grid *g, const float *const *tables, int npick, const float *obstime,
                  const float *pwt, const char *phase, int *isgood, float *resmin, float *tps, oracle_loc_out *out) {
    if (!g || g->nx < 2 || g->ny < 2 || g->nz < 2 || npick < 1 || g->ndiv < 1 || g->ndiv2 < 1) return -1;
    pass_t P;
    memset(&P, 0, sizeof P);
    P.g = g;
    P.tab = tables;
    P.npick = npick;
    P.obs = obstime;
    P.pwt = pwt;
    P.ph = phase;
    P.good = isgood;
    P.tp = (double *)malloc(sizeof(double) * (size_t)npick);
    P.res = (float *)malloc(sizeof(float) * (size_t)npick);
    P.resmin = resmin;
    P.tps = tps;
    memset(resmin, 0, sizeof(float) * (size_t)npick);
    memset(tps, 0, sizeof(float) * (size_t)npick);
    /* every pick counts in the first pass's variance, flagged or not (sphfdloc.c:683) */
    int ngood = npick, retried = 0, passes = 0;
    double e[3];
    for (;;) {
        P.ngood = ngood;
        P.stdmin = INFINITY;   /* (float)DBL_MAX */
        P.avrmin = 0.0f;
        search(&P, out->node, e);
        passes++;
        /* drop outliers (sphfdloc.c:1080-1096) */
        int redo = 0;
        ngood = 0;
        for (int p = 0; p < npick; p++)
            if (isgood[p]) {
                if (within(g, resmin[p], P.avrmin, tps[p]))
                    ngood++;
                else {
                    isgood[p] = 0;
                    redo = 1;
                }
            }
        if (ngood < g->nthres) break;
        if (redo) continue;
        /* once, bring back picks the new hypocentre explains (1109-1126) */
        if (retried < 1) {
            retried++;
            for (int p = 0; p < npick; p++)
                if (!isgood[p] && within(g, resmin[p], P.avrmin, tps[p])) {
                    ngood++;
                    isgood[p] = 1;
                    redo = 1;
                }
            if (redo) continue;
        }
        break;
    }
    out->status = ngood < g->nthres ? 0 : P.stdmin <= g->stdmax ? 1 : 2;
    out->ngood = ngood;
    out->passes = passes;
    out->stdmin = P.stdmin;
    out->avrmin = P.avrmin;
    out->exm = e[0];
    out->eym = e[1];
    out->ezm = e[2];
    free(P.tp);
    free(P.res);
    return 0;
}
