/*
 * fdtomoc_amd.h -- C ABI of the MI355X (gfx950) implementation of FDtomoC's
 * spherical finite-difference eikonal solver `sphfd`, with the steps either
 * side of it on the device: c2f's coarse-to-fine model interpolation and
 * sphfdloc's grid-search location over the resident tables.
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
    FDGPU_EBUSY = -10,      /* serving: arena/staging/queue full (retry after releasing
                               finished requests), or a call that allocates or frees
                               device memory made while the request loop runs */
    FDGPU_ETIMEOUT = -11,   /* serving: request not finished within the wait's timeout */
    FDGPU_ESCRATCH = -12,   /* fdgpu_solve_noalloc: the table's device state exceeds the
                               library's fixed 8 MiB array */
    FDGPU_ESPACE = -13      /* fdgpu_rays_trace: the derivative arrays are too small (the
                               results state how many entries each ray needs) */
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
/* Replaces the model's velocities (e.g. after an inversion step), in the
 * model's own device buffer (no allocation). */
int fdgpu_model_update(fdgpu_ctx *ctx, int model, const float *velocity);
int fdgpu_model_release(fdgpu_ctx *ctx, int model);

/* ------------------------------------------------------ coarse models (c2f)
 *
 * The fine velocity grids of an inversion's coarse model, computed on the
 * device and stored as resident models: the interpolation of c2f
 * (src/c2f.c:549-575, find_vel 673-727), which produces the P and S velocity
 * files sphfd reads.  Results are bit-identical to the files c2f writes
 * (payload after the 232-byte header: P then S, x fastest). */
typedef struct {
    int nxc, nyc, nzc;                    /* coarse nodes per axis, >= 2 each (c2f spec nxc nyc nzc) */
    const int *igridx, *igridy, *igridz;  /* fine intervals per coarse interval: nxc-1, nyc-1, nzc-1
                                             entries, each >= 1 (spec igridx/igridy/igridz) */
    double x0, y0, z0, h;                 /* spec origin and fine spacing */
    const float *velocity;                /* 2*nxc*nyc*nzc coarse velocities, P then S, x fastest
                                             (the coarse model file's payload) */
} fdgpu_coarse;

/* Fine dimensions of a coarse model: 1 + the sum of igrid per axis (src/c2f.c:262-277). */
int fdgpu_coarse_dims(const fdgpu_coarse *coarse, int *nx, int *ny, int *nz);
/* Creates two models (P and S) from a coarse model, replacing c2f's run and
 * the file round trip to sphfd (src/c2f.c main, run.sh:26-28).  fine == NULL
 * or 2*nx*ny*nz floats: receives the fine velocities c2f would write. */
int fdgpu_models_from_coarse(fdgpu_ctx *ctx, const fdgpu_coarse *coarse, int *model_p, int *model_s, float *fine);
/* Replaces two existing models (of the coarse model's fine dimensions) with
 * the interpolation of an updated coarse model (one inversion iteration). */
int fdgpu_models_update_from_coarse(fdgpu_ctx *ctx, int model_p, int model_s, const fdgpu_coarse *coarse,
                                    float *fine);

/* Grid-search location over resident tables (sphfdloc, src/sphfdloc.c).
 * The grid is in sphfdloc's units after its set-up (sphfdloc.c:382-393):
 * x0 longitude and y0 geocentric colatitude of the origin in radians, z0
 * depth in km, df/dq the longitude/colatitude spacings in radians, h in km;
 * the remaining fields are the program's parameters (sphfdloc.c:111-128,
 * spec keys of the same names). */
typedef struct {
    int nx, ny, nz;
    double x0, y0, z0, df, dq, h;
    int kmin;                   /* first layer of the node search (2) */
    int ndiv, ndiv2;            /* divisions of the two refinements (20, 20; at most 512) */
    int ivs;                    /* 0: an S pick's time is vpvs times its table (1) */
    float vpvs;                 /* 1.78 */
    int nthres;                 /* fewest usable picks (8) */
    float resthres, resthrep;   /* outlier limits, seconds and percent (0.5, 5.0) */
    float stdmax;               /* largest accepted deviation (15.0) */
} fdgpu_loc_params;

/* One event: its picks (the lines of its event file, sphfdloc.c:612-676) and,
 * on return, its location.  isgood is read and updated by the outlier passes;
 * res and tcalc receive each pick's residual and computed time at the
 * hypocentre (resmin and tps of sphfdloc.c:799-801; zero where no pass found
 * a minimum). */
typedef struct {
    int npick;                  /* 1..800 (maxobs) */
    const int *table;           /* pick p's resident table (fdgpu_solve), of the grid's dimensions */
    const float *obstime;       /* arrival time minus the event's reference time, seconds */
    const float *pwt;           /* weight 1/rwt^2 (sphfdloc.c:650) */
    const char *phase;          /* 'P' or 'S' */
    int *isgood;                /* 1 usable, 0 flagged ('*'); updated */
    float *res;                 /* out: npick residuals */
    float *tcalc;               /* out: npick computed times */
    int status;                 /* out: 1 located; 2 enough picks, deviation above stdmax; 0 too few picks */
    int ngood;                  /* out: usable picks after the outlier passes */
    int passes;                 /* out: searches run */
    int node[3];                /* out: node of the last node search */
    float stdmin, avrmin;       /* out: deviation and mean weighted residual at the hypocentre */
    double exm, eym, ezm;       /* out: hypocentre (longitude, colatitude in radians; depth in km) */
} fdgpu_loc_event;

/* sphfdloc's parameter defaults (sphfdloc.c:111-128); grid fields zero. */
void fdgpu_loc_params_defaults(fdgpu_loc_params *p);
/* The defaults plus the search grid as sphfdloc sets it up from the spec's
 * grid (sphfdloc.c:382-393): origin longitude/latitude in degrees (latitude
 * geodetic), depth z0 and spacing h in km, dq/df in degrees (0: from h).
 * z0r receives the radius of the grid's origin (km), for fdgpu_loc_geographic. */
int fdgpu_loc_params_from_grid(int nx, int ny, int nz, double x0_deg, double y0_deg, double z0, double h,
                               double dq_deg, double df_deg, fdgpu_loc_params *p, double *z0r);
/* A hypocentre (exm, eym, ezm of fdgpu_loc_event) as sphfdloc prints it
 * (sphfdloc.c:1131-1133): longitude and geodetic latitude in degrees and the
 * ellipsoidal depth in km (WGS84, geographic_method.c glathinv). */
int fdgpu_loc_geographic(double exm, double eym, double ezm, double z0, double z0r, double *lat_deg,
                         double *lon_deg, double *depth);
/* Locates n events (replaces sphfdloc's per-event loop, sphfdloc.c:752-1130,
 * and its table reads): each search pass runs on the device for every event
 * still being located; the outlier and retry decisions run on the host
 * between passes.  Grids below 2^30 nodes.  FDGPU_EINPUT for bad parameters, an unknown table or one
 * whose dimensions are not the grid's, or a phase other than P and S. */
int fdgpu_locate(fdgpu_ctx *ctx, const fdgpu_loc_params *p, fdgpu_loc_event *events, int n);
/* Measurements of the last fdgpu_locate (no reference counterpart): t[0]
 * search passes, t[1] device time of the passes (ms), t[2] of the node search
 * alone (ms), t[3] the table bytes the node searches read (4 bytes per table
 * a usable pick names, per searched node, per pass); n <= 4 values. */
int fdgpu_locate_timing(fdgpu_ctx *ctx, double *t, int n);

/* ------------------------------------------------ ray tracing (sphrayderv)
 *
 * The rays of sphrayderv (src/sphrayderv.c:1306-2250, lib/sphrayderv
 * dfind.c, pmin.c) through resident travel-time tables: from each event
 * back to its station, cell by cell down the gradient of the station's
 * table, with each segment's length spread trilinearly onto the coarse
 * inversion grid at its midpoint.  Replaces the per-pick loop of
 * sphrayderv.c:1069-2252 and its table reads (1163-1217); the caller keeps
 * the bookkeeping around it (weights, residual statistics, S-variable
 * offsets, the matrix files).  Results equal the reference's except where
 * the device's sine, cosine and arctangent of a ray position differ from the
 * host C library's in the last bit (derivatives and residuals are float
 * results of double steps; such a difference rarely reaches them).
 *
 * The grid in sphrayderv's units after its set-up (sphrayderv.c:527-545,
 * 814-826): x0 longitude and y0 geocentric colatitude of the origin in
 * radians, z0 depth in km, df/dq radians, h km; gx/gy/gz the coarse node
 * coordinates as the program accumulates them (floats). */
typedef struct {
    int nx, ny, nz;                 /* fine grid: every table's dimensions */
    double x0, y0, z0, df, dq, h;
    int nxc, nyc, nzc;              /* coarse nodes per axis, >= 2 */
    const float *gx, *gy, *gz;      /* nxc, nyc, nzc coarse coordinates (host arrays) */
    int ido1d;                      /* 1: derivatives per coarse layer (sphrayderv.c:1567-1571) */
    int max_segments;               /* segments before a ray is given up (0: 8*(nx+ny+nz));
                                       the reference has no limit */
} fdgpu_ray_grid;

/* One pick's ray: its station-phase table and the event and station in the
 * grid's units -- the event's ex, ey, ez (sphrayderv.c:949-951) and the
 * station's stlon, stlat (float values) and stz (sphrayderv.c:874-876). */
typedef struct {
    int table;
    double ex, ey, ez;
    double xs, ys, zs;
} fdgpu_ray;

enum {
    FDGPU_RAY_OK = 1,          /* reached the station's cell: derivatives valid */
    FDGPU_RAY_UNREACHED = 0,   /* a corner of the event's cell holds no time (> 1e9; sphrayderv.c:1340-1350) */
    FDGPU_RAY_LEFT = -1,       /* left the medium (sphrayderv.c:2248-2251) */
    FDGPU_RAY_EVENT_OUT = -2,  /* event outside the grid (sphrayderv.c:958-962) */
    FDGPU_RAY_LIMIT = -3,      /* more than max_segments segments */
    FDGPU_RAY_COARSE = -4,     /* a segment midpoint beyond the coarse grid (the reference asserts) */
    FDGPU_RAY_EDGE = -5        /* a cell whose corners lie outside the table (the reference reads
                                  memory beyond the table there) */
};

typedef struct {
    int status;                 /* FDGPU_RAY_* */
    int nseg;                   /* segments traced */
    int ndu;                    /* nonzero derivatives (FDGPU_RAY_OK only) */
    int first_in_cube;          /* 1: the first segment ran inside the station cube; ai/az are NaN
                                   (the reference keeps its previous ray's values there) */
    double tev;                 /* trilinear table time at the event (dt before corrections, 1385-1390) */
    float dtdx, dtdy, dtdz;     /* its derivatives (1404-1445) */
    float ai, az;               /* incidence and azimuth at the source, degrees (1865-1873) */
    float rdev;                 /* mean deviation from the source-station plane (1592) */
    long long offset;           /* first derivative of this ray in du_index/du_value */
} fdgpu_ray_result;

/* Traces n rays.  The derivatives of the rays that reached their station go
 * to du_index (coarse variable, P numbering: the caller adds nxc*nyc*nzc
 * for S picks as sphrayderv.c:1653-1656 does) and du_value (path length in
 * km, before weighting), each ray's in index order, rays in input order,
 * at results[r].offset.  When the total exceeds du_cap the call returns
 * FDGPU_ESPACE with every result filled and nothing written to du_*: call
 * again with room for the sum of ndu.  FDGPU_EINPUT for an unknown table or
 * one whose dimensions are not the grid's, or bad grid values. */
int fdgpu_rays_trace(fdgpu_ctx *ctx, const fdgpu_ray_grid *grid, const fdgpu_ray *rays, int n,
                     fdgpu_ray_result *results, int *du_index, float *du_value, long long du_cap);

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

/* One table into a caller's device buffer, with no device allocation (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same solve as fdgpu_solve, bit for bit, by one
 * workgroup: the input is resident model d->model (created beforehand), the
 * output d->par.nx*ny*nz floats written to `d_time` (device memory of the
 * caller; for srctype 2/3 the preset is copied into it first).  Every other
 * piece of device state -- the table's prepared arrays, counters, geometry
 * tables and face scratch -- lives in one fixed 8 MiB array of the library,
 * one per GPU, held until the process ends: a table whose largest face needs
 * more (about 130,000 nodes) is refused with FDGPU_ESCRATCH.  Calls on one GPU
 * share that array and must not overlap (they race); the reverse-pass skip
 * is off here (it needs a per-node array), results are unchanged.  stats:
 * NULL or nstats counters as fdgpu_table_stats returns them.  Synchronous;
 * FDGPU_EBUSY while serving is on. */
int fdgpu_solve_noalloc(fdgpu_ctx *ctx, const fdgpu_table_desc *d, float *d_time, int *stats, int nstats);

/* Re-solves n resident tables in place against their models' current
 * velocities (after fdgpu_model_update: an inversion iteration), reusing the
 * tables' device memory -- sphfd's run over the same par files with updated
 * velocity files (src/sphfd.c:200-257), without the per-table allocations,
 * uploads and releases of a new fdgpu_solve.  Allocates nothing once an
 * earlier solve or re-solve of the context has sized the per-workgroup
 * scratch for tables this large.  Tables of srctype 2/3 (whose initial grid
 * the solve overwrote) and loaded tables are refused (FDGPU_EINPUT); a table
 * with background writes in flight returns FDGPU_EBUSY.  Same results as
 * releasing and solving them again. */
int fdgpu_resolve(fdgpu_ctx *ctx, const int *table_ids, int n, int mode);

/* Device time of the last fdgpu_solve's (or coarse-model call's) kernel
 * launches, in milliseconds, measured with HIP events on the launch stream. */
double fdgpu_last_kernel_ms(fdgpu_ctx *ctx);

/* Per-face schedule knob (no reference counterpart; results do not depend on
 * it).  Reverse-pass faces first try the CHANGE schedule (evaluate the whole
 * face at once, then re-evaluate what changed); when changes cascade through a
 * face it is handed to the LEVEL schedule (each node once, in dependency
 * order) after `change_budget` evaluations per face node (0: never).  Default
 * 8; fdtomoc_amd/autotune.py picks it per workload by timing a sample. */
int fdgpu_set_schedule(fdgpu_ctx *ctx, int change_budget);
int fdgpu_get_schedule(fdgpu_ctx *ctx, int *change_budget);
/* Workgroup shape of batched launches (no reference counterpart; results do
 * not depend on it): 512 lanes and one workgroup (table) per CU, or 256 lanes
 * and two per CU.  0 (default): 512 while the tables fit one per CU, else 256. */
int fdgpu_set_workgroup(fdgpu_ctx *ctx, int lanes);
/* Wavefront passes (no reference counterpart; results do not depend on it).
 * A pass whose bit is set runs as one dataflow wavefront over the whole
 * table: a node is evaluated as soon as every neighbour that precedes it in
 * the reference's face-by-face, time-sorted order is final, so the faces of a
 * pass overlap instead of running one after another.  Bit 0: the first pass
 * (box growth); bit 1: reverse passes.  Default 0: face by face, the
 * faster on most measured workloads (fdtomoc_amd/autotune.py times both on
 * the caller's tables).  Tables of srctype 2/3 and tables whose wavefront
 * cannot finish in the slot's state run face by face. */
int fdgpu_set_wave(fdgpu_ctx *ctx, int flags);
int fdgpu_get_wave(fdgpu_ctx *ctx, int *flags);

/* Teams (no reference counterpart; results are identical for every choice):
 * in a batched solve with wavefront passes, the largest tables have their
 * wavefront passes shared by several workgroups on as many CUs (2..16)
 * instead of one.  size 0 (default): planned per table -- each table gets
 * the smallest team expected to finish within a common batch time, the
 * shortest whose work fills the device's workgroups at `share_pct` % (0:
 * 100); size 2..16: that size for the leading tables of the largest-first
 * queue with at least `share_pct` % (0: 200) of the mean nodes per
 * workgroup of a full device; size 1: never. */
int fdgpu_set_team(fdgpu_ctx *ctx, int size, int share_pct);
int fdgpu_get_team(fdgpu_ctx *ctx, int *size, int *share_pct);

/* Counters of the context (no reference counterpart): c[0] tables solved a
 * second time, face by face, because one of the sweep kernel's own
 * consistency checks failed on the first solve (the second solve's results
 * are kept; a second failure is FDGPU_EKERNEL); c[1] tables solved a second
 * time because a wavefront pass could not finish (list capacity, a request
 * that never resolved, a team member that waited too long); n <= 2 values. */
int fdgpu_context_counters(fdgpu_ctx *ctx, long long *c, int n);

/* Copies a resident table (nx*ny*nz floats) to host memory. */
int fdgpu_table_read(fdgpu_ctx *ctx, int table, float *host, unsigned long long count);
/* Writes a resident table as the reference does (src/sphfd.c:5702-5803):
 * `parse` decimation, 232-byte header derived from `vel_header` (the velocity
 * file's 232 bytes as the reference holds them, i.e. after its swab 1|3 of
 * the numeric fields, sphfd.c:668-682; or NULL for the header the reference
 * builds for a headerless velocity file, sphfd.c:615-656), optional byte swap
 * (par.swab 2|3). */
int fdgpu_table_write(fdgpu_ctx *ctx, int table, const char *path, const void *vel_header);
/* Loads a travel-time table file into a new resident table, read as
 * FDtomoC's consumers read one (sphrayderv.c:1184-1217, sphfdloc's
 * read_timefiles): a file whose header's first token is "HEAD" has its
 * nx*ny*nz native floats after the 232-byte header, any other file from byte
 * 0.  For fdgpu_locate and fdgpu_table_read; fdgpu_table_write refuses such a
 * table (no solve inputs behind it).  FDGPU_EIO: missing or short file. */
int fdgpu_table_load(fdgpu_ctx *ctx, const char *path, int nx, int ny, int nz, int *table);
/* Writes n resident tables as fdgpu_table_write does, in the background: a
 * worker thread copies them to pinned host memory on a stream of its own and
 * writes the files, so the copies and file writes overlap the caller's next
 * fdgpu_solve (sphfd.c:5702-5803 off the critical path; no reference
 * counterpart).  vel_headers: NULL, or n pointers each NULL or 232 bytes as
 * fdgpu_table_write takes them; paths and headers are copied.  A table with
 * writes in flight cannot be released (FDGPU_EBUSY). */
int fdgpu_tables_write_async(fdgpu_ctx *ctx, const int *tables, const char *const *paths,
                             const void *const *vel_headers, int n);
/* Waits for every background write; returns the first error since the last
 * wait (FDGPU_EIO for a file that could not be written), or 0. */
int fdgpu_write_wait(fdgpu_ctx *ctx);
/* Sweep counters of a table: passes, restarts, LEVEL faces, CHANGE faces,
 * LEVEL rounds, CHANGE rounds, fallbacks, node evaluations, deferrals, error
 * bits, last restart wall, then the first pass's head-wave counts (7). */
int fdgpu_table_stats(fdgpu_ctx *ctx, int table, int *stats, int nstats);
int fdgpu_table_release(fdgpu_ctx *ctx, int table);
/* Device pointer of a resident table (for zero-copy consumers on the GPU). */
int fdgpu_table_device_ptr(fdgpu_ctx *ctx, int table, void **dptr);

/* ------------------------------------------------------------------ serving
 *
 * A device-resident request loop with continuous batching (no reference
 * counterpart: the reference solves one table per call).  While requests are
 * outstanding a persistent kernel runs with one workgroup per CU; each
 * workgroup claims the next request as soon as it finishes its previous one,
 * so a request that arrives while others run starts on the first free
 * workgroup and overlaps with them.  Results are bit-identical to
 * fdgpu_solve's.  Admission is by memory: each request takes a block of a
 * fixed HBM arena (its time grid and stencil tables, plus its slowness if it
 * carries its own velocities) and of a pinned staging area for its inputs;
 * when either is full submit returns FDGPU_EBUSY.
 *
 * While serving is on, calls that allocate or free device memory
 * (fdgpu_solve, fdgpu_model_*, fdgpu_table_release) return FDGPU_EBUSY:
 * create the resident models first. */
typedef struct {
    int max_nx, max_ny, max_nz;       /* largest table to admit (sizes per-workgroup scratch) */
    unsigned long long arena_bytes;   /* HBM arena for admitted requests (0: all the budget leaves) */
    unsigned long long staging_bytes; /* pinned staging for request inputs (default 1 GiB) */
    int queue_capacity;               /* requests in flight (default 4096) */
    int workgroups;                   /* server workgroups (0: one per CU) */
} fdgpu_serve_config;

void fdgpu_serve_config_defaults(fdgpu_serve_config *cfg);
int fdgpu_serve_start(fdgpu_ctx *ctx, const fdgpu_serve_config *cfg);
/* Admits one table.  velocity == NULL: the table is a window of resident model
 * d->model at (ox, oy, oz).  Otherwise `velocity` holds the request's own
 * par.nx*ny*nz velocities (e.g. the model after an inversion update; origin
 * must be 0); it is copied before the call returns. */
int fdgpu_serve_submit(fdgpu_ctx *ctx, const fdgpu_table_desc *d, const float *velocity, int *request);
/* Admits one table whose velocities are phase `phase` (0 = P, 1 = S) of a
 * coarse model (fdgpu_coarse, as an inversion iteration emits it): the
 * workgroup that takes the request interpolates its slowness as c2f does
 * (src/c2f.c:549-575, 673-727) and then solves it; the coarse model's fine
 * grid must be the table's (origin 0).  No reference counterpart: it replaces
 * c2f's run, its files and sphfd's re-read of them for one table.  Such a
 * request skips no reverse-pass evaluation (no velocity range is known before
 * the interpolation); results are identical either way. */
int fdgpu_serve_submit_coarse(fdgpu_ctx *ctx, const fdgpu_table_desc *d, const fdgpu_coarse *coarse, int phase,
                              int *request);
int fdgpu_serve_poll(fdgpu_ctx *ctx, int request, int *done);
/* Waits until the request is finished (timeout_s < 0: no limit; FDGPU_ETIMEOUT). */
int fdgpu_serve_wait(fdgpu_ctx *ctx, int request, double timeout_s);
int fdgpu_serve_read(fdgpu_ctx *ctx, int request, float *host, unsigned long long count);
int fdgpu_serve_stats(fdgpu_ctx *ctx, int request, int *stats, int nstats);
/* Writes a finished request's table as fdgpu_table_write does. */
int fdgpu_serve_write(fdgpu_ctx *ctx, int request, const char *path, const void *vel_header);
/* HBM arena bytes one request with these sizes takes, for sizing
 * cfg->arena_bytes: own_velocity, it carries its own velocities; coarse (or
 * NULL), it carries this coarse model (fdgpu_serve_submit_coarse; implies
 * own_velocity). */
int fdgpu_serve_request_bytes(const fdgpu_par *par, int own_velocity, const fdgpu_coarse *coarse,
                              unsigned long long *bytes);
/* Returns a finished request's arena block. */
int fdgpu_serve_release(fdgpu_ctx *ctx, int request);
/* info[0..5]: serving on, requests outstanding, arena bytes used, arena bytes,
 * server launches so far, server workgroups. */
int fdgpu_serve_info(fdgpu_ctx *ctx, unsigned long long *info, int ninfo);
/* Finishes every admitted request, stops the loop and frees its memory
 * (results of unreleased requests are dropped). */
int fdgpu_serve_stop(fdgpu_ctx *ctx);

/* Drop-in of src/sphfd.c:260: av[1] is one parlist line ("par=STA.par").
 * Solves on device 0 and writes output_path/timefile.  Returns 0 or a
 * negative FDGPU_E* code (the reference aborts instead).  Thread-safe:
 * calls made concurrently (as the reference's OpenMP loop makes them,
 * sphfd.c:247-255) are gathered into one batched launch per velocity input
 * -- the first caller to find no launch running waits until no other call
 * arrived for 1 ms (at most 5 ms) and solves every call queued by then. */
int sphfd(int ac, char **av, char *output_path);
/* Counters of the in-process entry point (no reference counterpart):
 * info[0] sphfd() calls, info[1] launches they were solved in. */
int fdgpu_sphfd_info(unsigned long long *info, int ninfo);

/* The `sphfd` executable's body (src/sphfd.c:200-257) with all tables of the
 * parlist solved in one batched launch per velocity file.  Reads the spec path
 * from `spec_path` (the executable passes what it read from stdin). */
int fdgpu_run_spec(const char *spec_path);

#ifdef __cplusplus
}
#endif
#endif
