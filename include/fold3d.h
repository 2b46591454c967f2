/*
 * fold3d.h — C-ABI of the MI355X (gfx950) parallel marching method (PMM)
 * solver for the 3D anisotropic eikonal equation.
 *
 * Replaces the single-device solve of the reference
 *   Stencil-OpenACC/AN_ISO_1_OpenACC.c:1150-1241 (driver: device copy, sweep loop)
 *   Stencil-OpenACC/AN_ISO_1_OpenACC.c:180-1092  (six plane-step kernels)
 * behind plain C types: pointers, sizes, doubles.  No HIP or torch types appear
 * in any signature; streams travel as `void *` (a hipStream_t, NULL = the
 * legacy default stream).
 *
 * Grid layout (reference Fold3dPMM.h:27-30): T[i + (nx+2)*(j + (ny+2)*k)],
 * i fastest, nodes 1..n per axis, one ghost layer on each side that the solver
 * reads but never writes.  T is signed: the sign marks the side of the initial
 * surface and is preserved; |T| >= FOLD3D_MAXT means "not reached".
 *
 * Conventions: every entry point returns FOLD3D_OK (0) or a negative
 * FOLD3D_E* code, never exits.  The caller owns every buffer passed in; a
 * handle owns its scratch memory.  A handle is used by one host thread at a
 * time; separate handles are independent.
 */
#ifndef FOLD3D_H
#define FOLD3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOLD3D_MAXT (10e15 * 1.0) /* reference Fold3dPMM.h:6 */

enum fold3d_status {
  FOLD3D_OK = 0,
  FOLD3D_E_INVALID = -1,     /* bad argument (size, pointer, option) */
  FOLD3D_E_HIP = -2,         /* a HIP runtime call failed (see fold3d_last_error) */
  FOLD3D_E_NOMEM = -3,       /* device allocation failed */
  FOLD3D_E_UNSUPPORTED = -4, /* valid request this library does not implement */
  FOLD3D_E_NOT_CONVERGED = -5, /* max_iters reached before tol (result still written) */
  FOLD3D_E_IO = -6             /* a file could not be opened, read or written */
};

enum fold3d_precision { FOLD3D_F64 = 0, FOLD3D_F32 = 1 };

enum fold3d_model_kind {
  /* Reference model: F*|grad T| + a.grad T = 1, a flips sign where T < 0
   * (AN_ISO_1_OpenACC.c:206-212). */
  FOLD3D_MODEL_FOLD = 0,
  /* Tilted elliptic model: sqrt(grad T^T M grad T) = 1/v(x), M symmetric
   * positive definite, constant or per node. */
  FOLD3D_MODEL_ELLIP = 1
};

/* Where the array T lives. */
enum fold3d_memory { FOLD3D_DEVICE = 0, FOLD3D_HOST = 1 };

typedef struct fold3d_model {
  int kind;         /* enum fold3d_model_kind */
  /* FOLD3D_MODEL_FOLD */
  double F;         /* reference: 1.1  (Fold3dPMM.h:52) */
  double a[3];      /* reference: (-0.35, 0.4, 0.7) (Fold3dPMM.h:49-51) */
  /* FOLD3D_MODEL_ELLIP, constant part: M = [[m0 m3 m4][m3 m1 m5][m4 m5 m2]] */
  double M[6];
  double speed;     /* used when speed_field == NULL */
  /* FOLD3D_MODEL_ELLIP, optional per-node fields on the padded grid, same
   * precision as T, device pointers (NULL = use the constant):              */
  const void *speed_field;            /* v(x) > 0 */
  const void *theta, *phi, *ra, *rb;  /* tilt angles (rad) and axis ratios:
                                         M = R diag(1, ra, rb)^2 R^T,
                                         R = Rz(phi) Ry(theta); all four or none */
} fold3d_model;

/* Sweep order: the reference's (AN_ISO_1_OpenACC.c:1195-1236). */
enum fold3d_direction {
  FOLD3D_X_DOWN = 0, /* +x: reads plane i, writes i+1, i = 1..n-1 */
  FOLD3D_X_UP = 1,   /* -x */
  FOLD3D_K_UP = 2,   /* +z */
  FOLD3D_K_DOWN = 3, /* -z */
  FOLD3D_Y_UP = 4,   /* -y (reference name) */
  FOLD3D_Y_DOWN = 5  /* +y */
};

typedef struct fold3d_opts {
  int precision;    /* enum fold3d_precision of T (and of per-node fields) */
  int memory;       /* enum fold3d_memory of T */
  int max_iters;    /* outer iterations of 6 sweeps; reference: 8 */
  double tol;       /* > 0: stop after the first iteration with max|dT| < tol;
                       0: run exactly max_iters (the reference's behaviour) */
  void *stream;     /* hipStream_t; NULL = default stream */
  int engine;       /* 0 = auto; see enum fold3d_engine */
} fold3d_opts;

enum fold3d_engine {
  FOLD3D_ENGINE_AUTO = 0,    /* the fastest engine available for the model;
                                for the fold model SPLIT, and on a solve that
                                may run more than 24 iterations LOCKED from
                                iteration 24 on (same bits; locking pays only
                                once most rows have settled) */
  FOLD3D_ENGINE_PLANE = 1,   /* one launch per plane, node update written as the
                                reference's local solvers (fold model) */
  FOLD3D_ENGINE_HOISTED = 2, /* one launch per plane, per-direction invariants
                                hoisted out of the node update (same bits) */
  FOLD3D_ENGINE_TILED = 3,   /* hoisted node update, source tile staged in LDS,
                                each sweep pair run in a layout whose march axis
                                is slowest (needs a scratch copy of T, owned by
                                the handle) */
  FOLD3D_ENGINE_SPLIT = 4,   /* as TILED, with each node's candidate units spread
                                over several wavefronts whose minima meet in LDS
                                (same bits; what AUTO runs for the fold model in
                                its first 24 iterations; for the elliptic model
                                the split or metric engine) */
  FOLD3D_ENGINE_LOCKED = 5   /* fold model: SPLIT plus exact locking — from the
                                iteration set by fold3d_set_lock_start (default
                                8, 0-based), a block of target rows whose source
                                patch has not changed since its previous update
                                in the same direction is not recomputed (its
                                update provably returns the value it has).  Same
                                bits as SPLIT; stats->updates_computed tells how
                                many cell-updates were evaluated */
};

typedef struct fold3d_stats {
  int iterations;        /* outer iterations run */
  double max_dT;         /* max |T_new - T_old| over the last iteration (tol > 0 only) */
  double cell_updates;   /* iterations * 6 * (n-1) * n^2 for cubes */
  double sweep_ms;       /* device time of the sweep loop (HIP events) */
  double updates_computed; /* cell-updates evaluated: cell_updates, or fewer with
                              FOLD3D_ENGINE_LOCKED (rows skipped by locking) */
  double bytes_moved;    /* algorithmic HBM bytes of the plane steps: (2 + P) *
                            sizeof(T) per evaluated update (read and write the
                            target, read P per-node fields: speed_field, and
                            theta, phi, ra, rb when given); layout moves are not
                            counted.  0 from the adjoint entry points */
  double nodes_held;     /* fold3d_adjoint: reached nodes whose value no candidate
                            reproduces (within 1e-12 relative in FP64, 1e-6 in
                            FP32), held fixed: they pass no gradient upstream.
                            The closed-form start around a source is such a set;
                            more means lost gradient paths.  0 from fold3d_solve */
} fold3d_stats;

typedef struct fold3d_handle fold3d_handle;

/* Library ---------------------------------------------------------------- */
const char *fold3d_version(void);
const char *fold3d_last_error(void);     /* thread-local message of the last failure */

int fold3d_create(fold3d_handle **out, int device);
int fold3d_destroy(fold3d_handle *h);

void fold3d_default_opts(fold3d_opts *o);           /* F64, device, 8 iters, tol 0 */
void fold3d_reference_model(fold3d_model *m);       /* F = 1.1, a = (-0.35, 0.4, 0.7) */

/* Whole solve, in place (replaces AN_ISO_1_OpenACC.c:1185-1241).
 * dx, dy, dz: node spacings.  Returns FOLD3D_E_NOT_CONVERGED (with T and
 * stats written) when tol > 0 and max_iters ran out first.  During the solve
 * T also serves as one of the two buffers the field moves between layouts
 * through, so after any other error its contents are unspecified.
 * T must hold no NaN (the reference's unreached value is FOLD3D_MAXT); the
 * engines take minima of magnitudes in one instruction, which equals the
 * reference's compare-and-select only for numbers.
 * Precision, fold model: FP64 gives the reference's bits with every engine;
 * FP32 gives, with every engine, the bits of the reference's arithmetic run
 * in float.  On the reference workload (8 iterations, up to 512^3) the FP32
 * field keeps every sign and the reached set, and against FP64 its relative
 * error is below 1e-2 everywhere, above 1e-3 at no more than 0.5 % of the
 * reached nodes, above 1e-4 at 10 %, above 1e-5 at 25 % (where FP32 rounding
 * flips one of the reference's discontinuous acceptance tests, a node takes
 * another candidate).  Elliptic model: FP32 within about 3e-6 of FP64. */
int fold3d_solve(fold3d_handle *h, void *T, int nx, int ny, int nz,
                 double dx, double dy, double dz, const fold3d_model *model,
                 const fold3d_opts *opts, fold3d_stats *stats);

/* fold3d_solve without a handle and without allocating: no hipMalloc, hipFree
 * or other allocating call, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it).  No
 * counterpart in the reference.  Fold model, device-resident T, FP64 or FP32;
 * engine AUTO or SPLIT (others are FOLD3D_E_UNSUPPORTED).  Same bits as
 * fold3d_solve.  Its scratch is a fixed 32 KiB per GPU in the library's static
 * device memory (the six directions' constant tables and the tolerance stop's
 * max|dT| words, whatever the grid), held until the process ends; grids too
 * large for a device are FOLD3D_E_NOMEM before any launch.  Every call on one
 * GPU shares that scratch: calls that overlap on one GPU (several streams or
 * threads) race.  All six directions run on the caller's field in the
 * reference layout (no second field), so the x-sweeps stage strided rows and
 * the solve is slower than fold3d_solve's.  With tol > 0 the call waits for
 * the stream once per iteration (the stop test), so only tol == 0 suits graph
 * capture.  stats->sweep_ms is 0 (the call creates no events). */
int fold3d_solve_noalloc(void *T, int nx, int ny, int nz, double dx, double dy, double dz,
                         const fold3d_model *model, const fold3d_opts *opts,
                         fold3d_stats *stats);

/* Batched elliptic-model solves: many independent instances, each with its
 * own grid, spacings, metric and speed (constant or per node), in one stream
 * of launches (one launch per sweep direction and plane step serves every
 * instance that has that plane; persistent blocks pull rows of 64 nodes from a
 * device queue).  The launch sequence of one iteration is captured in a
 * hipGraph and replayed (opts->engine == FOLD3D_ENGINE_PLANE launches it
 * directly instead).  All arrays are device arrays in opts->precision, on the
 * padded layout of their own grid.  An instance stops being updated once an
 * iteration changes none of its nodes by opts->tol or more.  Outputs (each
 * may be NULL, arrays of length `count`): iterations run and last max|dT| per
 * instance, and the device time of the whole batch. */
typedef struct fold3d_instance {
  void *T;
  int n[3];
  double h[3];
  double M[6];                        /* used when theta..rb are NULL */
  double speed;                       /* used when speed_field is NULL */
  const void *speed_field;
  const void *theta, *phi, *ra, *rb;  /* all four or none */
} fold3d_instance;

int fold3d_solve_batched(fold3d_handle *h, int count, const fold3d_instance *inst,
                         const fold3d_opts *opts, int *iterations_out,
                         double *max_dt_out, double *ms_out);

/* Adjoint state of a converged elliptic-model field T (device, opts->precision)
 * for the misfit J = 1/2 sum_r (T[rec_idx[r]] - T_obs[r])^2: `rec_res` holds
 * the residuals T[rec] - T_obs at the `nrec` receiver nodes (device arrays of
 * linear node indices and values).  Adds dJ/d(slowness) per node to
 * grad_slowness and, with per-node tilt fields, dJ/d(theta) to grad_theta
 * (may be NULL); both device arrays on the padded layout.  The linear upwind
 * system of T is solved by reverse-ordered sweeps until no adjoint value
 * changes by more than opts->tol (opts->max_iters at most).  Slowness = 1/v,
 * so the model needs a speed field or a constant speed; nodes no candidate
 * reproduces (the closed-form start around a source) are held fixed. */
int fold3d_adjoint(fold3d_handle *h, const void *T, int nx, int ny, int nz,
                   double dx, double dy, double dz, const fold3d_model *model,
                   int nrec, const int *rec_idx, const void *rec_res,
                   const fold3d_opts *opts, void *grad_slowness, void *grad_theta,
                   fold3d_stats *stats);

/* fold3d_adjoint for `count` sources on one grid and model at once (e.g. the
 * sources of a tomography survey): every plane launch of the reverse sweeps
 * serves all sources still running, so small grids fill the GPU.  Each source
 * has its converged field, receivers, residuals and stop rule (`tol`, as
 * opts->tol of fold3d_adjoint; opts->tol is not used) and stops on its own.
 * The gradients of all sources are added to grad_slowness / grad_theta in
 * source order, with the same bits as `count` calls of fold3d_adjoint in that
 * order.  Scratch: the handle holds 80 B per node per source in FP32 (132 B
 * in FP64).  Outputs (may be NULL): sweeps run, last max change and reached
 * nodes held fixed (see fold3d_stats.nodes_held) per source (arrays of
 * length `count`), device time.  FOLD3D_E_NOT_CONVERGED if any
 * source ran out of opts->max_iters sweeps (gradients written). */
typedef struct fold3d_adjoint_source {
  const void *T;          /* converged field of the source (device, opts->precision) */
  int nrec;               /* receivers: linear node indices (device int32) and */
  const int *rec_idx;     /*   residuals T[rec] - T_obs (device, opts->precision) */
  const void *rec_res;
  double tol;             /* stop when no adjoint value changes by more than tol */
} fold3d_adjoint_source;

int fold3d_adjoint_batched(fold3d_handle *h, int count, const fold3d_adjoint_source *src,
                           int nx, int ny, int nz, double dx, double dy, double dz,
                           const fold3d_model *model, const fold3d_opts *opts,
                           void *grad_slowness, void *grad_theta, int *iterations_out,
                           double *max_d_out, long long *held_out, double *ms_out);

/* Tuning of the split engines, per model kind: how many wavefronts share one
 * row of 64 nodes (fold model: 2, 4 or 8, default 4; elliptic model with a
 * constant metric: 1, 2 or 4, default 2; 0 restores the default).
 * fold3d_autotune() times every choice on `iterations` iterations of a
 * scratch copy of T (T itself is not modified), keeps the fastest on the
 * handle for that model kind and returns its sweep time. */
int fold3d_set_tuning(fold3d_handle *h, int model_kind, int groups);
/* First iteration (0-based) that FOLD3D_ENGINE_LOCKED runs locked; the ones
 * before run as FOLD3D_ENGINE_SPLIT.  Default 8 (-1 restores it): in the
 * reference's 8 iterations most rows still change, and a locked launch costs
 * more per computed row; locking pays on the way to the fixed point. */
int fold3d_set_lock_start(fold3d_handle *h, int iteration);
int fold3d_get_lock_start(fold3d_handle *h, int *iteration);
int fold3d_get_tuning(fold3d_handle *h, int model_kind, int *groups);
int fold3d_autotune(fold3d_handle *h, const void *T, int nx, int ny, int nz,
                    double dx, double dy, double dz, const fold3d_model *model,
                    const fold3d_opts *opts, int iterations, float *ms_out);

/* Start field of an elliptic-model solve from a point source at `src` (in
 * the grid's coordinates: node (i,j,k) sits at ((i-1)dx, (j-1)dy, (k-1)dz)):
 * T = sqrt(d^T M^-1 d) / speed, d = node - src, at the interior nodes within
 * `radius` cells of the source in every axis (1.6 in the tests and
 * benchmarks), FOLD3D_MAXT elsewhere and on the ghost layer.  With per-node
 * fields, pass the metric and speed at the node nearest the source.  T is a
 * device array in `precision`; asynchronous on `stream`. */
int fold3d_point_source(fold3d_handle *h, void *T, int nx, int ny, int nz, double dx,
                        double dy, double dz, const double src[3], const double M[6],
                        double speed, double radius, int precision, void *stream);

/* The reference's input field (ImplicitInitialiser, Fold3dPMM.c:129-389)
 * built on the device into a device FP64 array T on the padded grid, with the
 * same bits as the host initializer: the device evaluates the implicit
 * surface everywhere, and the host's libm evaluates it again at the nodes
 * whose exact value matters (near the surface: ends of edges where the sign
 * changes, and nodes where the sign is in doubt); *n_exact (may be NULL)
 * receives their number.  Records the grid for the reference-signature plane
 * steps like ImplicitInitialiser.  Synchronous on `stream`. */
int fold3d_initialiser_device(fold3d_handle *h, double *T, int nx, int ny, int nz, double dx,
                              double dy, double dz, void *stream, long long *n_exact);

/* Tuning of the batched solve: waves per row of 64 nodes (1, 2, 4; default
 * 1), persistent blocks per compute unit in each launch (1..64; default 64)
 * and how a block finds its rows: 1..64 = that many rows per pull from the
 * launch's device queue, 0 = queue with an automatic pull size, -1 = static
 * (block b takes rows b, b + grid, ...; the default).  0 arguments of
 * set restore the defaults of groups and blocks.  fold3d_autotune_batched()
 * times a set of these on `iterations` iterations of scratch copies of the
 * batch's fields (which are not modified), keeps the fastest on the handle
 * and returns its device time. */
int fold3d_set_batch_tuning(fold3d_handle *h, int groups, int blocks_per_cu, int rows_per_pull);
int fold3d_get_batch_tuning(fold3d_handle *h, int *groups, int *blocks_per_cu,
                            int *rows_per_pull);
int fold3d_autotune_batched(fold3d_handle *h, int count, const fold3d_instance *inst,
                            const fold3d_opts *opts, int iterations, float *ms_out);
/* Sub-batches the batched solve runs at once (1..8; 0 restores the default,
 * 3).  The instances are cut into that many bands of about equal work, the
 * largest instances first, and each band is solved as a batch of its own on
 * a stream of its own, from a host thread of its own, all ordered after
 * opts->stream; fold3d_solve_batched returns when all are done.  Every
 * instance gets the bits it gets in one batch; the bands keep the GPU busy
 * while another waits on its convergence check or runs a plane step that few
 * of its instances have.  No counterpart in the reference (no batched
 * solve there). */
int fold3d_set_batch_streams(fold3d_handle *h, int streams);
int fold3d_get_batch_streams(fold3d_handle *h, int *streams);

/* One plane step of one direction on a device FP64 field, fold model, with
 * explicit local-frame arguments exactly as the reference kernels take them
 * (AN_ISO_1_OpenACC.c:180-183; the CUDA twin Stencil-CUDA/AN_ISO_2_GPU.cu:151-155
 * uses F, ay, ax, az as passed, as this does).  Asynchronous on `stream`. */
int fold3d_plane_step(fold3d_handle *h, int direction, double *T, int nx,
                      int ny, int nz, int plane, double dzz, double dxx,
                      double dyy, double dxy, double dxz, double dyz, double F,
                      double ay, double ax, double az, void *stream);

/* The same per direction, named after the reference's six kernels
 * (function_x_down AN_ISO_1_OpenACC.c:180, function_x_up :334, function_y_up
 * :485, function_y_down :636, function_k_up :790, function_k_down :944), with
 * the grid, the handle and the stream added: fold3d_function_x_down(h, T, nx,
 * ny, nz, plane, ...) == fold3d_plane_step(h, FOLD3D_X_DOWN, T, nx, ny, nz,
 * plane, ...).  A driver loop over several grids or streams calls these. */
#define FOLD3D_PLANE_STEP_ARGS                                                  \
  fold3d_handle *h, double *T, int nx, int ny, int nz, int plane, double dzz,  \
      double dxx, double dyy, double dxy, double dxz, double dyz, double F,    \
      double ay, double ax, double az, void *stream
int fold3d_function_x_down(FOLD3D_PLANE_STEP_ARGS);
int fold3d_function_x_up(FOLD3D_PLANE_STEP_ARGS);
int fold3d_function_k_up(FOLD3D_PLANE_STEP_ARGS);
int fold3d_function_k_down(FOLD3D_PLANE_STEP_ARGS);
int fold3d_function_y_up(FOLD3D_PLANE_STEP_ARGS);
int fold3d_function_y_down(FOLD3D_PLANE_STEP_ARGS);

/* Measured ceilings of the current device, for reporting next to the
 * published peaks: sustained FP64 and FP32 FMA rate (FLOP/s) and the
 * bandwidth of a plain device-to-device copy (bytes read + written per s).
 * Any output pointer may be NULL.  Allocates 4 GiB briefly for the copy. */
int fold3d_measure_peaks(void *stream, double *fp64_flops, double *fp32_flops,
                         double *copy_bytes_per_s);

/* Host files of a padded FP64 field (host pointers).
 * fold3d_write_text / fold3d_read_text: the reference's solution dump
 *   (AN_ISO_1_OpenACC.c:1244-1260): "%f\n" per node, ghosts included, i
 *   outermost and k innermost; the reader takes any whitespace-separated
 *   numbers in that order and fails on too few or too many.
 * fold3d_write_vtk: legacy VTK STRUCTURED_POINTS of the interior nodes, origin
 *   0 at node (1,1,1), spacing dx, dy, dz; binary != 0 writes big-endian
 *   doubles (replaces the header-only Write_VTK_Structured_Grid,
 *   AN_ISO_1_OpenACC.c:1098-1135). */
int fold3d_write_text(const char *path, const double *T, int nx, int ny, int nz);
int fold3d_read_text(const char *path, double *T, int nx, int ny, int nz);
int fold3d_write_vtk(const char *path, const double *T, int nx, int ny, int nz, double dx,
                     double dy, double dz, int binary);

/* Device memory helpers for C callers that have no HIP headers
 * (stand in for acc_malloc / acc_memcpy_* at AN_ISO_1_OpenACC.c:1185-1186,1241). */
int fold3d_device_alloc(void **ptr, size_t bytes);
int fold3d_device_free(void *ptr);
int fold3d_copy_to_device(void *dst, const void *src, size_t bytes);
int fold3d_copy_to_host(void *dst, const void *src, size_t bytes);
int fold3d_synchronize(void *stream);

#ifdef __cplusplus
}
#endif
#endif
