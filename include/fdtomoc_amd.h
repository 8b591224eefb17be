/*
 * fdtomoc_amd.h -- C ABI of the MI355X (gfx950) implementation of FDtomoC's
 * spherical finite-difference eikonal solver `sphfd`.
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the reference's contract
 *       int sphfd(int ac, char **av, char *output_path)   (src/sphfd.c:194, 260)
 *     and the `sphfd` executable (stdin spec file, src/sphfd.c:200-257), which
 *     read the same par/spec files and write the same travel-time tables.
 *
 *  2. A batched/served GPU API: velocity models resident in HBM, many station
 *     tables solved in one persistent launch, tables kept resident between
 *     calls and read back or written in the reference layout on demand.
 *
 * No C++ types cross this boundary.  All functions return 0 on success or a
 * negative FDGPU_E* code; fdgpu_strerror() describes a code.
 */
#ifndef FDTOMOC_AMD_H
#define FDTOMOC_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FDGPU_OK = 0,
    FDGPU_EHIP = -1,        /* HIP runtime error (no device, launch failure...) */
    FDGPU_ENOMEM = -2,      /* device allocation failed or HBM budget exceeded */
    FDGPU_ESOURCE = -3,     /* source outside the model (src/sphfd.c:533-538) */
    FDGPU_ESRCTYPE = -4,    /* srctype not 1, 2 or 3 (src/sphfd.c:516-519) */
    FDGPU_EINPUT = -5,      /* bad argument or missing input data */
    FDGPU_EIO = -6,         /* file open/read/write failure */
    FDGPU_EPAR = -7,        /* required par key missing (src/sphfd.c:6526, exit 100) */
    FDGPU_ESTOP = -8,       /* STOP in par input (src/sphfd.c:6462-6465, exit 101) */
    FDGPU_EKERNEL = -9,     /* the sweep kernel reported an internal error */
    FDGPU_EBUSY = -10       /* serving queue full */
};

/* Par-file values of one table (keys of src/sphfd.c:383-530), before the
 * reference's conversions.  Angles in degrees (dq/df 0 = derive from h),
 * depths in km positive down. */
typedef struct {
    int nx, ny, nz;
    double h, dq, df, x0, y0, z0;
    int iminus, iplus, jminus, jplus, kminus, kplus;
    int floatsrc, srctype, srcwall;
    double fxs, fys, fzs;
    int xs, ys, zs;
    int ncube, reverse, headpref, parse, fill;
    float maxoff;
    int invh, invq, invf, invx0, invy0, invz0;
    int swab;
} fdgpu_par;

/* Fills `p` with the reference's defaults (src/sphfd.c:266-367). */
void fdgpu_par_defaults(fdgpu_par *p);

typedef struct fdgpu_ctx fdgpu_ctx;

/* Opens device `device` with an HBM budget in bytes (0 = 90% of free memory). */
int fdgpu_open(int device, unsigned long long hbm_budget_bytes, fdgpu_ctx **ctx);
int fdgpu_close(fdgpu_ctx *ctx);
const char *fdgpu_strerror(int code);

/* A velocity model on the device: nx*ny*nz float32 velocities, x fastest
 * (src/sphfd.c:151-152).  Stored as slowness 1/v rounded to float as
 * src/sphfd.c:689-690 computes it.  `model` receives an id. */
int fdgpu_model_create(fdgpu_ctx *ctx, int nx, int ny, int nz, const float *velocity, int *model);
/* Replaces the model's velocities (e.g. after an inversion step). */
int fdgpu_model_update(fdgpu_ctx *ctx, int model, const float *velocity);
int fdgpu_model_release(fdgpu_ctx *ctx, int model);

/* One table request.  Table node (0,0,0) sits at model node (ox, oy, oz); the
 * table spans par.nx*par.ny*par.nz nodes inside the model. */
typedef struct {
    fdgpu_par par;
    int model;
    int ox, oy, oz;
    const float *preset;   /* srctype 2: six wall planes as src/sphfd.c:829-888 reads
                              them; srctype 3: old table values; else NULL */
    const float *boxvals;  /* fill = 1: box-file values (src/sphfd.c:718-719); else NULL */
} fdgpu_table_desc;

enum { FDGPU_BATCHED = 0, FDGPU_PLAIN = 1 };

/* Solves n tables.  FDGPU_BATCHED: one persistent launch whose workgroups pull
 * tables from a queue (largest first).  FDGPU_PLAIN: one launch per table, in
 * order.  The tables stay resident; ids are written to table_ids[0..n). */
int fdgpu_solve(fdgpu_ctx *ctx, const fdgpu_table_desc *tables, int n, int mode, int *table_ids);

/* Device time of the last fdgpu_solve's kernel launches, in milliseconds,
 * measured with HIP events on the launch stream. */
double fdgpu_last_kernel_ms(fdgpu_ctx *ctx);

/* Copies a resident table (nx*ny*nz floats) to host memory. */
int fdgpu_table_read(fdgpu_ctx *ctx, int table, float *host, unsigned long long count);
/* Writes a resident table as the reference does (src/sphfd.c:5702-5803):
 * `parse` decimation, 232-byte header derived from `vel_header` (232 bytes of
 * the velocity file, or NULL for the reference's headerless fallback),
 * optional byte swap (par.swab 2|3). */
int fdgpu_table_write(fdgpu_ctx *ctx, int table, const char *path, const void *vel_header);
/* Sweep counters of a table: passes, restarts, LEVEL faces, CHANGE faces,
 * LEVEL rounds, CHANGE rounds, fallbacks, node evaluations, deferrals, error
 * bits, last restart wall, then the first pass's head-wave counts (7). */
int fdgpu_table_stats(fdgpu_ctx *ctx, int table, int *stats, int nstats);
int fdgpu_table_release(fdgpu_ctx *ctx, int table);
/* Device pointer of a resident table (for zero-copy consumers on the GPU). */
int fdgpu_table_device_ptr(fdgpu_ctx *ctx, int table, void **dptr);

/* Drop-in of src/sphfd.c:260: av[1] is one parlist line ("par=STA.par").
 * Solves on device 0 and writes output_path/timefile.  Returns 0 or a
 * negative FDGPU_E* code (the reference aborts instead). */
int sphfd(int ac, char **av, char *output_path);

/* The `sphfd` executable's body (src/sphfd.c:200-257) with all tables of the
 * parlist solved in one batched launch per velocity file.  Reads the spec path
 * from `spec_path` (the executable passes what it read from stdin). */
int fdgpu_run_spec(const char *spec_path);

#ifdef __cplusplus
}
#endif
#endif
