/*
 * sphfd_oracle.h -- CPU restatement of FDtomoC's spherical finite-difference
 * eikonal solver (src/sphfd.c), used ONLY as a test checker and as the CPU
 * baseline of bench.py.  Nothing in the fdtomoc_amd product links it.
 *
 * Parity status (see the tests under tests/):
 *   - whole solve: pinned bit-for-bit against the reference-shipped fixture
 *     verification/cmod/M001.ptimes and against the reference compiled from
 *     its own sources (oracle/_ref/sphfd_ref) on seeded models;
 *   - stencils: pinned by closed-form known-answer tests (planar fields);
 *   - geometry (glath): pinned by an independent WGS84 formulation.
 */
#ifndef SPHFD_ORACLE_H
#define SPHFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Raw par-file values, before the conversions of sphfd.c:415-509.  Field
 * names follow the par keys (sphfd.c:383-413, 460-530).  dq/df are degrees
 * (0 = derive from h), x0/y0 geographic degrees, z0/fzs depth in km. */
typedef struct {
    int nx, ny, nz;
    double h, dq, df, x0, y0, z0;
    int iminus, iplus, jminus, jplus, kminus, kplus;
    int floatsrc, srctype, srcwall;
    double fxs, fys, fzs;      /* floatsrc = 1 */
    int xs, ys, zs;            /* floatsrc = 0 */
    int ncube, reverse, headpref, parse, fill;
    float maxoff;
    int invh, invq, invf, invx0, invy0, invz0;
} oracle_par;

/* Converted geometry (sphfd.c:415-509). */
typedef struct {
    double h, dq, df, x0, y0, z0r;   /* radians, y0 = geocentric colatitude */
    int xs, ys, zs;                  /* source node */
    double fxss, fyss, fzss;         /* header source: lon rad, colat rad, radius km */
    int maxrad;
} oracle_geom;

typedef struct {
    int passes;          /* iterations of the outer (reverse) loop */
    int restarts;        /* reverse passes started (a "RESTART" message each) */
    long long faces;     /* faces swept */
    long long visits;    /* node visits (trial sequences evaluated) */
    long long stores;    /* improved nodes written */
    long long stale_uninit;  /* 2D trials that read never-set t[0]/t[4] */
    long long carry_reads;   /* 2D trials whose t[0]/t[4] came from an earlier node */
    int last_srcwall;
    int headw_first[7];  /* head-wave counters after the first pass */
} oracle_stats;

/* p_ geometry per sphfd.c:415-509 (glath at 5836-5893).  Returns 0, or -1 if
 * the source is outside the model (sphfd.c:533-538). */
int oracle_geometry(const oracle_par *p, oracle_geom *g);

/* Full solve of one table.  slow: nx*ny*nz slownesses (1/v rounded to float,
 * sphfd.c:689-690).  time: output grid (overwritten).  extra: for srctype 2
 * the six wall planes (6*nwall floats as read at sphfd.c:829-888), for
 * srctype 3 the old time grid, for fill=1 the box values; NULL otherwise.
 * Returns 0 on success. */
int oracle_solve(const oracle_par *p, const oracle_geom *g, const float *slow,
                 float *time, const float *extra, oracle_stats *st);

/* The four stencils (sphfd.c:5906-6184), exported for known-answer tests.
 * They take the same argument layout as the reference functions. */
double oracle_fdsph3d(const double t[8], const double s[8], const double r[8],
                      const double q[8], const int g[8], const int n[8],
                      const int m[8], double h, double dq, double df);
double oracle_fdsphne(const double t[8], const double s[8], double d01,
                      double d12, double d14, double d25, double d45);
double oracle_fdsph2d(const double t[8], const double s[8], double d01,
                      double d02, double d13, double d23);
double oracle_fdsphnf(const double t[8], const double s[8], double d02,
                      double d12, double d25);
double oracle_glath(double hlat, double h, double *r);

#ifdef __cplusplus
}
#endif
#endif
