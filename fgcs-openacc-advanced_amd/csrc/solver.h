// solver.h — internal interfaces between the C-ABI layer and the engines.
#pragma once

#include "common.h"
#include "fold_fast.h"
#include "fold_local.h"

namespace fold3d {

// One plane step's addressing: node (a, b) of the plane, a, b in 1..na/nb,
// sits at base + a*Sa + b*Sb; the target is +sm away, the local x^/y^
// neighbours are +-su / +-sv away.
struct PlaneGeom {
  long long base, Sa, Sb;
  long long sm, su, sv;
  int na, nb;
  int ua, ub, va, vb;  // x^ and y^ as (a, b) offsets within a tile
};

PlaneGeom plane_geom(const Grid &G, int dir, int plane);

template <typename R>
int fold_sweeps_plane_engine(hipStream_t st, R *T, const Grid &G,
                             const fold3d_model &m, const double h[3],
                             int max_iters, double tol,
                             unsigned long long *d_max, int *iters_out,
                             double *max_dt_out);

template <typename R>
int ellip_sweeps_plane_engine(hipStream_t st, R *T, const Grid &G,
                              const fold3d_model &m, const double h[3],
                              int max_iters, double tol,
                              unsigned long long *d_max, int *iters_out,
                              double *max_dt_out);

// Constant metric: layouts + LDS rows + units spread over `groups` (1, 2, 4)
// waves.  Y = a T-sized scratch field; VX, VY = the speed field in the x- and
// y-sweep layouts (unused without a speed field).
template <typename R>
int ellip_sweeps_split_engine(hipStream_t st, R *T, R *Y, R *VX, R *VY, const Grid &G,
                              const fold3d_model &m, const double h[3], int max_iters,
                              double tol, int groups, unsigned long long *d_max,
                              int *iters_out, double *max_dt_out);

// Per-node metric (tilt and axis-ratio fields): slowness and the lattice
// products of N = M^-1 hoisted into kMetricFields fields in the x- and in the y-sweep layout once per
// solve, layouts + LDS rows + units over `groups` (1, 2, 4) waves.  `scratch`
// holds ellip_metric_scratch_fields() fields of G.count elements (a T copy
// and the two sets of metric fields).
constexpr int kMetricFields = 7;
size_t ellip_metric_scratch_fields();
template <typename R>
int ellip_sweeps_metric_engine(hipStream_t st, R *T, R *scratch, const Grid &G,
                               const fold3d_model &m, const double h[3], int max_iters,
                               double tol, int groups, unsigned long long *d_max,
                               int *iters_out, double *max_dt_out);

void sym3_inverse(const double M[6], double N[6]);

template <typename R>
void build_fold_fast(int dir, const fold3d_model &m, const double h[3], FoldFast<R> &out);

template <typename R>
int fold_sweeps_fast_engine(hipStream_t st, R *T, const Grid &G, const FoldFast<R> *d_consts,
                            int max_iters, double tol, unsigned long long *d_max,
                            int *iters_out, double *max_dt_out);

// Hoisted constants + LDS-staged tiles + per-direction layouts.  `Y` is a
// scratch field of G.count elements.
// `groups`: 0 = one thread per node (fold_tile_kernel); 2, 4 or 8 = the
// node's eight candidate units spread over that many waves (fold_split_kernel).
// `stamps` (line_stamp_words(G) ints, or null): exact locking of rows whose
// source patch has not changed since their last update in the same direction;
// `updates_computed` (device, 256 counters zeroed by the caller) counts the
// cell-updates the locked iterations computed; locking starts at iteration
// `lock_from` (0-based; the iterations before it run unlocked and are not
// counted).
template <typename R>
int fold_sweeps_tiled_engine(hipStream_t st, R *T, R *Y, const Grid &G,
                             const FoldFast<R> *d_consts, int max_iters, double tol, int groups,
                             unsigned long long *d_max, int *iters_out, double *max_dt_out,
                             int *stamps = nullptr, unsigned long long *updates_computed = nullptr,
                             int lock_from = 0);

// Locking stamps: one int per 64-node chunk of every x-line and every y-line
// of the padded grid.
inline int line_chunks(int n) { return (n + 2 + 63) / 64; }
inline size_t line_stamp_count(const Grid &G, bool ylines) {
  return ylines ? (size_t)(G.n[0] + 2) * (G.n[2] + 2) * line_chunks(G.n[1])
                : (size_t)(G.n[1] + 2) * (G.n[2] + 2) * line_chunks(G.n[0]);
}
inline size_t line_stamp_words(const Grid &G) {
  return line_stamp_count(G, false) + line_stamp_count(G, true);
}

// The split kernel (`groups` = 2, 4 or 8 waves per row) on T in the reference
// layout for all six directions: no scratch field.  d_consts: the six
// directions' tables; d_max: kMaxSlots words (tol > 0).
template <typename R>
int fold_sweeps_inplace_engine(hipStream_t st, R *T, const Grid &G, const FoldFast<R> *d_consts,
                               int max_iters, double tol, int groups, unsigned long long *d_max,
                               int *iters_out, double *max_dt_out);

// Batched elliptic solves (batch.hip).  `blocks` = persistent grid per launch,
// `groups` = waves per row (1, 2, 4), `rows_per_pull` = rows taken from the
// queue at a time (0 = automatic); `scratch` holds batch_scratch_bytes(): per
// instance a T copy and the metric fields in the x- and y-sweep layouts.
template <typename R>
int batch_solve(hipStream_t st, int count, const fold3d_instance *inst, int max_iters,
                double tol, int blocks, int groups, int rows_per_pull, int use_graph,
                void *scratch, int *iters_out, double *max_dt_out, double *ms_out);
size_t batch_scratch_bytes(int count, const fold3d_instance *inst, int precision);

// Adjoint state and gradients of `nsrc` sources on one grid and model
// (adjoint.hip); `scratch` holds adjoint_scratch_bytes(G, precision, nsrc)
// bytes.  iters_out / maxd_out / held_out (reached nodes held fixed): per source.
struct AdjointSource {
  const void *T;        // converged forward field (device)
  int nrec;
  const int *rec_idx;   // receiver node indices (device)
  const void *rec_res;  // residuals at the receivers (device)
  double tol;           // stop rule: max change of an adjoint value in a pass
};
template <typename R>
int adjoint_solve(hipStream_t st, int nsrc, const AdjointSource *srcs, const Grid &G,
                  const fold3d_model &m, const double h[3], int max_iters, R *grad_s,
                  R *grad_th, void *scratch, int *iters_out, double *maxd_out,
                  long long *held_out);
size_t adjoint_scratch_bytes(const Grid &G, int precision, int nsrc);

// The reference's input field built on the device, same bits as the host
// initializer (gpu_init.hip); n_exact_out = nodes whose IF the host evaluated.
int device_initialiser(hipStream_t st, double *T, int nx, int ny, int nz, double dx, double dy,
                       double dz, long long *n_exact_out);

// Start field of an elliptic solve from a point source (source_init.hip).
int point_source_start(hipStream_t st, void *T, int precision, int nx, int ny, int nz,
                       const double h[3], const double src[3], const double N[6], double v,
                       double radius);

// Copy a padded field from one axis order to another (layout.hip).
template <typename R>
void permute_layout(hipStream_t st, const R *in, const Grid &Gin, R *out, const Grid &Gout);

// Where the field of a running solve lives.  The x-, z- and y-sweeps each run
// in their own layout (X, Z, Y), so an iteration moves the field X -> Z -> Y,
// and the move Y -> X opens the next iteration: three moves per iteration, and
// one back to the reference layout at the end.  Every move writes the other
// of two buffers (the caller's field and a scratch field of the same size).
// The buffer the field ends in follows from the number of moves; the engines
// move it back to Z once more after the first iteration when that makes a
// fixed iteration count end in the caller's buffer, and finish() copies it
// there otherwise (a solve stopped by its tolerance).
template <typename R>
struct FieldMover {
  hipStream_t st;
  R *buf[2];
  const Grid *lay;
  int cur = 0;

  FieldMover(hipStream_t s, R *T, R *scratch, const Grid &GZ) : st(s), buf{T, scratch}, lay(&GZ) {}
  R *field() const { return buf[cur]; }
  void to(const Grid &G) {
    if (lay == &G) return;
    permute_layout<R>(st, buf[cur], *lay, buf[1 - cur], G);
    cur = 1 - cur;
    lay = &G;
  }
  hipError_t finish(const Grid &GZ) {
    to(GZ);
    if (cur == 0) return hipSuccess;
    cur = 0;
    return hipMemcpyAsync(buf[0], buf[1], sizeof(R) * GZ.count, hipMemcpyDeviceToDevice, st);
  }
};

int fold_plane_step_f64(hipStream_t st, double *T, const Grid &G, int dir,
                        int plane, const FoldFrame<double> &f);

}  // namespace fold3d

// Grid shared by the reference-signature entry points (host_init.c).
extern "C" void fold3d_get_grid(int *nx, int *ny, int *nz);
