// capi.hip — the extern "C" surface of libfold3d.so (include/fold3d.h and the
// device half of include/Fold3dPMM.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include <cmath>
#include <cstring>

#include "../../include/Fold3dPMM.h"
#include "common.h"
#include "solver.h"

using namespace fold3d;

struct fold3d_handle {
  int device;
  unsigned long long *d_max;  // max |dT| of the running iteration (kMaxSlots words)
  hipEvent_t ev0, ev1;
  void *d_consts;             // per-direction constants of the running solve
  void *h_consts;             // pinned staging for them
  hipEvent_t consts_copied;   // staging may be rewritten once this completes
  void *d_scratch;            // a second field for the layout changes
  size_t scratch_bytes;
  int split_groups;           // waves per node row of the fold split kernel (2, 4, 8)
  int ellip_groups;           // waves per node row of the elliptic split kernel (1, 2, 4)
  int batch_groups;           // batched solve: waves per row (1, 2, 4)
  int batch_blocks_per_cu;    // batched solve: persistent blocks per CU per launch
  int batch_rows_per_pull;    // batched solve: rows per queue pull (0 = automatic)
  int batch_streams;          // batched solve: sub-batches run at once (1..kMaxBatchStreams)
  int lock_from;              // locked engine: first locked iteration (0-based)
};

#define F3D_CHECK_DEV(p, what)                                                             \
  F3D_CHECK((p) == nullptr || is_device_ptr(p), FOLD3D_E_INVALID,                          \
            "%s must be device memory (got a host or unknown pointer)", what)

// True when the device can address p: device, managed or pinned host memory.
// A plain (unregistered) host pointer passed where a device array is required
// is rejected instead of being dereferenced by a kernel.
static bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory reports an error: clear it
    return false;
  }
  return a.type != hipMemoryTypeUnregistered;
}

static int check_model_fields(const fold3d_model *m) {
  F3D_CHECK_DEV(m->speed_field, "model speed_field");
  F3D_CHECK_DEV(m->theta, "model theta");
  F3D_CHECK_DEV(m->phi, "model phi");
  F3D_CHECK_DEV(m->ra, "model ra");
  F3D_CHECK_DEV(m->rb, "model rb");
  return FOLD3D_OK;
}

// Device memory released when it goes out of scope (error paths included).
struct DeviceBuffer {
  void *p = nullptr;
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
};

// Grow the handle's scratch field to at least `bytes` (stream-ordered free of
// the old one is not available, so the device is synchronised first).
static int ensure_scratch(fold3d_handle *h, size_t bytes) {
  if (h->scratch_bytes >= bytes) return FOLD3D_OK;
  if (h->d_scratch) {
    F3D_HIP(hipDeviceSynchronize());
    F3D_HIP(hipFree(h->d_scratch));
    h->d_scratch = nullptr;
    h->scratch_bytes = 0;
  }
  F3D_HIP(hipMalloc(&h->d_scratch, bytes));
  h->scratch_bytes = bytes;
  return FOLD3D_OK;
}

static constexpr size_t kConstBytes = 6 * sizeof(FoldFast<double>);

// First locked iteration of FOLD3D_ENGINE_LOCKED by default: the reference's
// 8 iterations run unlocked (most rows still change in them, and a locked
// launch costs more per computed row), the iterations after them locked.
static constexpr int kDefaultLockFrom = 8;
// FOLD3D_ENGINE_AUTO on the fold model: the split engine, locked from this
// iteration on when the solve may run longer.  Measured at 512^3 FP64: while
// 80-98 % of the rows still change (the first 24 iterations) a locked
// iteration costs 3-4 % more than a split one; beyond that locking wins
// (1.17x over 80 iterations, 1.47x to the fixed point).  Same bits either way.
static constexpr int kAutoLockFrom = 24;
// Batched solve: sub-batches run at once by default, and at most.  On the
// 96-instance workload (sides 64..320, FP32 to 1e-6) 1, 2, 3, 4 and 6 bands
// ran at 117.5, 134.1, 135.5, 119.7 and 134.7 solves/s.
static constexpr int kDefaultBatchStreams = 3;
static constexpr int kMaxBatchStreams = 8;

// False when one FP64 field of the padded grid would exceed 2^56 bytes (far
// beyond any device, and below the point where the size arithmetic of the
// engines, 15 fields of it, could overflow 64 bits).
static bool grid_bytes_fit(long long nx, long long ny, long long nz) {
  const double bytes = 8.0 * (double)(nx + 2) * (double)(ny + 2) * (double)(nz + 2);
  return bytes < 0x1p56;
}

extern "C" {

const char *fold3d_version(void) { return "fold3d-mi355x 0.1 (gfx950)"; }

void fold3d_default_opts(fold3d_opts *o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->precision = FOLD3D_F64;
  o->memory = FOLD3D_DEVICE;
  o->max_iters = 8;  // AN_ISO_1_OpenACC.c:1187
  o->tol = 0.0;
  o->stream = nullptr;
  o->engine = FOLD3D_ENGINE_AUTO;
}

void fold3d_reference_model(fold3d_model *m) {
  if (!m) return;
  memset(m, 0, sizeof *m);
  m->kind = FOLD3D_MODEL_FOLD;
  m->F = 1.1;  // Fold3dPMM.h:49-52
  m->a[0] = -0.35;
  m->a[1] = 0.4;
  m->a[2] = 0.7;
  m->speed = 1.0;
}

int fold3d_create(fold3d_handle **out, int device) {
  F3D_CHECK(out != nullptr, FOLD3D_E_INVALID, "fold3d_create: out is NULL");
  int count = 0;
  F3D_HIP(hipGetDeviceCount(&count));
  F3D_CHECK(device >= 0 && device < count, FOLD3D_E_INVALID,
            "fold3d_create: device %d not in [0,%d)", device, count);
  F3D_HIP(hipSetDevice(device));
  fold3d_handle *h = new fold3d_handle();  // value-initialised: no resources yet
  h->device = device;
  h->split_groups = 4;
  h->ellip_groups = 2;
  h->batch_groups = 1;
  h->batch_blocks_per_cu = 64;
  h->batch_rows_per_pull = -1;
  h->batch_streams = kDefaultBatchStreams;
  h->lock_from = kDefaultLockFrom;
  hipError_t e = hipMalloc((void **)&h->d_max, sizeof(unsigned long long) * kMaxSlots);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->consts_copied, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&h->d_consts, kConstBytes);
  if (e == hipSuccess) e = hipHostMalloc(&h->h_consts, kConstBytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error("fold3d_create: %s", hipGetErrorString(e));
    fold3d_destroy(h);  // releases what was created
    return e == hipErrorOutOfMemory ? FOLD3D_E_NOMEM : FOLD3D_E_HIP;
  }
  *out = h;
  return FOLD3D_OK;
}

int fold3d_destroy(fold3d_handle *h) {
  if (!h) return FOLD3D_OK;
  (void)hipSetDevice(h->device);
  if (h->consts_copied) (void)hipEventSynchronize(h->consts_copied);
  if (h->d_scratch) (void)hipFree(h->d_scratch);
  if (h->d_consts) (void)hipFree(h->d_consts);
  if (h->h_consts) (void)hipHostFree(h->h_consts);
  if (h->consts_copied) (void)hipEventDestroy(h->consts_copied);
  if (h->d_max) (void)hipFree(h->d_max);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  delete h;
  return FOLD3D_OK;
}

int fold3d_solve(fold3d_handle *h, void *T, int nx, int ny, int nz, double dx,
                 double dy, double dz, const fold3d_model *model,
                 const fold3d_opts *opts, fold3d_stats *stats) {
  F3D_CHECK(h && T && model, FOLD3D_E_INVALID, "fold3d_solve: NULL handle, T or model");
  F3D_CHECK(nx >= 1 && ny >= 1 && nz >= 1, FOLD3D_E_INVALID,
            "fold3d_solve: grid %dx%dx%d too small (need >= 1 per axis)", nx, ny, nz);
  F3D_CHECK(dx > 0 && dy > 0 && dz > 0 && std::isfinite(dx) && std::isfinite(dy) &&
                std::isfinite(dz),
            FOLD3D_E_INVALID, "fold3d_solve: spacings must be positive and finite");
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  F3D_CHECK(o.max_iters >= 1, FOLD3D_E_INVALID, "fold3d_solve: max_iters must be >= 1");
  F3D_CHECK(o.tol >= 0 && std::isfinite(o.tol), FOLD3D_E_INVALID, "fold3d_solve: bad tol");
  F3D_CHECK(o.precision == FOLD3D_F64 || o.precision == FOLD3D_F32, FOLD3D_E_INVALID,
            "fold3d_solve: bad precision %d", o.precision);
  F3D_CHECK(o.memory == FOLD3D_DEVICE || o.memory == FOLD3D_HOST, FOLD3D_E_INVALID,
            "fold3d_solve: bad memory kind %d", o.memory);
  F3D_CHECK(o.engine >= FOLD3D_ENGINE_AUTO && o.engine <= FOLD3D_ENGINE_LOCKED, FOLD3D_E_INVALID,
            "fold3d_solve: bad engine %d", o.engine);
  F3D_CHECK(o.engine != FOLD3D_ENGINE_LOCKED || model->kind == FOLD3D_MODEL_FOLD,
            FOLD3D_E_UNSUPPORTED, "fold3d_solve: locking is implemented for the fold model only");
  F3D_CHECK(model->kind == FOLD3D_MODEL_FOLD || model->kind == FOLD3D_MODEL_ELLIP,
            FOLD3D_E_INVALID, "fold3d_solve: unknown model kind %d", model->kind);
  F3D_CHECK(grid_bytes_fit(nx, ny, nz), FOLD3D_E_NOMEM,
            "fold3d_solve: a %dx%dx%d field does not fit in device memory", nx, ny, nz);
  if (model->kind == FOLD3D_MODEL_FOLD) {
    F3D_CHECK(std::isfinite(model->F) && model->F > 0, FOLD3D_E_INVALID,
              "fold3d_solve: F must be positive");
    // F|p| + a.p = 1 is an eikonal equation only for |a| < F (the wind must be
    // slower than the medium); the pruning bounds dzz/(F + az) and
    // dzz/(sqrt(F^2 - ay^2) + az) assume it too
    const double *a = model->a;
    F3D_CHECK(std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]) &&
                  a[0] * a[0] + a[1] * a[1] + a[2] * a[2] < model->F * model->F,
              FOLD3D_E_INVALID, "fold3d_solve: a must be finite with |a| < F");
  } else {
    const int nf = (model->theta != nullptr) + (model->phi != nullptr) +
                   (model->ra != nullptr) + (model->rb != nullptr);
    F3D_CHECK(nf == 0 || nf == 4, FOLD3D_E_INVALID,
              "fold3d_solve: give all four of theta, phi, ra, rb or none");
    F3D_CHECK(model->speed_field != nullptr || (std::isfinite(model->speed) && model->speed > 0),
              FOLD3D_E_INVALID, "fold3d_solve: speed must be positive");
    if (nf == 0) {
      const double *M = model->M;
      bool fin = true;
      for (int k = 0; k < 6; ++k) fin = fin && std::isfinite(M[k]);
      const double m1 = M[0], m2 = M[0] * M[1] - M[3] * M[3];
      const double m3 = M[0] * (M[1] * M[2] - M[5] * M[5]) - M[3] * (M[3] * M[2] - M[5] * M[4]) +
                        M[4] * (M[3] * M[5] - M[1] * M[4]);
      F3D_CHECK(fin && m1 > 0 && m2 > 0 && m3 > 0, FOLD3D_E_INVALID,
                "fold3d_solve: metric M is not symmetric positive definite");
    }
    F3D_CHECK(o.memory == FOLD3D_DEVICE || (model->speed_field == nullptr && nf == 0),
              FOLD3D_E_INVALID, "fold3d_solve: per-node fields need a device-resident T");
  }

  F3D_HIP(hipSetDevice(h->device));
  if (o.memory == FOLD3D_DEVICE) F3D_CHECK_DEV(T, "fold3d_solve: T");
  if (model->kind == FOLD3D_MODEL_ELLIP && check_model_fields(model) != FOLD3D_OK)
    return FOLD3D_E_INVALID;
  hipStream_t st = (hipStream_t)o.stream;
  const Grid G(nx, ny, nz);
  const size_t esz = o.precision == FOLD3D_F64 ? 8 : 4;
  const size_t bytes = (size_t)G.count * esz;
  void *dT = T;
  DeviceBuffer staging;  // host-resident T: its device copy, freed on every path
  if (o.memory == FOLD3D_HOST) {
    F3D_HIP(hipMalloc(&staging.p, bytes));
    dT = staging.p;
    F3D_HIP(hipMemcpyAsync(dT, T, bytes, hipMemcpyHostToDevice, st));
  }
  const double hsp[3] = {dx, dy, dz};
  int iters = 0;
  double maxdt = -1;
  F3D_HIP(hipEventRecord(h->ev0, st));
  int rc;
  unsigned long long *done_ctr = nullptr;  // cell-updates computed by locked iterations
  int locked_from = 0;                     // first locked iteration of this solve
  if (model->kind == FOLD3D_MODEL_FOLD && o.engine == FOLD3D_ENGINE_PLANE) {
    rc = o.precision == FOLD3D_F64
             ? fold_sweeps_plane_engine<double>(st, (double *)dT, G, *model, hsp, o.max_iters,
                                                o.tol, h->d_max, &iters, &maxdt)
             : fold_sweeps_plane_engine<float>(st, (float *)dT, G, *model, hsp, o.max_iters,
                                               o.tol, h->d_max, &iters, &maxdt);
  } else if (model->kind == FOLD3D_MODEL_FOLD) {
    // stage the six directions' constants (the previous solve's copy must be done)
    F3D_HIP(hipEventSynchronize(h->consts_copied));
    if (o.precision == FOLD3D_F64) {
      FoldFast<double> *c = (FoldFast<double> *)h->h_consts;
      for (int d = 0; d < 6; ++d) build_fold_fast<double>(d, *model, hsp, c[d]);
    } else {
      FoldFast<float> *c = (FoldFast<float> *)h->h_consts;
      for (int d = 0; d < 6; ++d) build_fold_fast<float>(d, *model, hsp, c[d]);
    }
    F3D_HIP(hipMemcpyAsync(h->d_consts, h->h_consts, kConstBytes, hipMemcpyHostToDevice, st));
    F3D_HIP(hipEventRecord(h->consts_copied, st));
    if (o.engine == FOLD3D_ENGINE_HOISTED) {
      rc = o.precision == FOLD3D_F64
               ? fold_sweeps_fast_engine<double>(st, (double *)dT, G,
                                                 (const FoldFast<double> *)h->d_consts,
                                                 o.max_iters, o.tol, h->d_max, &iters, &maxdt)
               : fold_sweeps_fast_engine<float>(st, (float *)dT, G,
                                                (const FoldFast<float> *)h->d_consts,
                                                o.max_iters, o.tol, h->d_max, &iters, &maxdt);
    } else {
      const bool lock = o.engine == FOLD3D_ENGINE_LOCKED ||
                        (o.engine == FOLD3D_ENGINE_AUTO && o.max_iters > kAutoLockFrom);
      const int lock_from = o.engine == FOLD3D_ENGINE_LOCKED ? h->lock_from : kAutoLockFrom;
      locked_from = lock_from;
      // scratch: T copy [, line-chunk stamps, computed-updates counter]
      const size_t tbytes = (bytes + 255) & ~(size_t)255;
      const size_t sbytes = (line_stamp_words(G) * sizeof(int) + 255) & ~(size_t)255;
      rc = ensure_scratch(h, lock ? tbytes + sbytes + 256 * 8 : bytes);
      if (rc != FOLD3D_OK) return rc;
      char *scr = (char *)h->d_scratch;
      int *stamps = lock ? (int *)(scr + tbytes) : nullptr;
      done_ctr = lock ? (unsigned long long *)(scr + tbytes + sbytes) : nullptr;
      if (lock) F3D_HIP(hipMemsetAsync(done_ctr, 0, 256 * sizeof(unsigned long long), st));
      const int groups = o.engine == FOLD3D_ENGINE_TILED ? 0 : h->split_groups;
      rc = o.precision == FOLD3D_F64
               ? fold_sweeps_tiled_engine<double>(st, (double *)dT, (double *)h->d_scratch, G,
                                                  (const FoldFast<double> *)h->d_consts,
                                                  o.max_iters, o.tol, groups, h->d_max, &iters,
                                                  &maxdt, stamps, done_ctr, lock_from)
               : fold_sweeps_tiled_engine<float>(st, (float *)dT, (float *)h->d_scratch, G,
                                                 (const FoldFast<float> *)h->d_consts,
                                                 o.max_iters, o.tol, groups, h->d_max, &iters,
                                                 &maxdt, stamps, done_ctr, lock_from);
    }
  } else if (model->theta == nullptr && o.engine != FOLD3D_ENGINE_PLANE &&
             o.engine != FOLD3D_ENGINE_HOISTED) {
    // constant metric: split engine; scratch = T copy (+ speed field in two layouts)
    const int nscr = model->speed_field ? 3 : 1;
    rc = ensure_scratch(h, nscr * bytes);
    if (rc != FOLD3D_OK) return rc;
    char *scr = (char *)h->d_scratch;
    const int groups = h->ellip_groups;
    rc = o.precision == FOLD3D_F64
             ? ellip_sweeps_split_engine<double>(
                   st, (double *)dT, (double *)scr, (double *)(scr + bytes),
                   (double *)(scr + 2 * bytes), G, *model, hsp, o.max_iters, o.tol, groups,
                   h->d_max, &iters, &maxdt)
             : ellip_sweeps_split_engine<float>(
                   st, (float *)dT, (float *)scr, (float *)(scr + bytes),
                   (float *)(scr + 2 * bytes), G, *model, hsp, o.max_iters, o.tol, groups,
                   h->d_max, &iters, &maxdt);
  } else if (model->theta != nullptr && o.engine != FOLD3D_ENGINE_PLANE &&
             o.engine != FOLD3D_ENGINE_HOISTED) {
    // per-node metric: metric engine; scratch = T copy + the metric fields in
    // the x- and y-sweep layouts
    rc = ensure_scratch(h, ellip_metric_scratch_fields() * bytes);
    if (rc != FOLD3D_OK) return rc;
    const int groups = h->ellip_groups;
    rc = o.precision == FOLD3D_F64
             ? ellip_sweeps_metric_engine<double>(st, (double *)dT, (double *)h->d_scratch, G,
                                                  *model, hsp, o.max_iters, o.tol, groups,
                                                  h->d_max, &iters, &maxdt)
             : ellip_sweeps_metric_engine<float>(st, (float *)dT, (float *)h->d_scratch, G,
                                                 *model, hsp, o.max_iters, o.tol, groups,
                                                 h->d_max, &iters, &maxdt);
  } else
    rc = o.precision == FOLD3D_F64
             ? ellip_sweeps_plane_engine<double>(st, (double *)dT, G, *model, hsp, o.max_iters,
                                                 o.tol, h->d_max, &iters, &maxdt)
             : ellip_sweeps_plane_engine<float>(st, (float *)dT, G, *model, hsp, o.max_iters,
                                                o.tol, h->d_max, &iters, &maxdt);
  F3D_HIP(hipEventRecord(h->ev1, st));
  if (rc != FOLD3D_OK) return rc;
  if (o.memory == FOLD3D_HOST) {
    F3D_HIP(hipMemcpyAsync(T, dT, bytes, hipMemcpyDeviceToHost, st));
    F3D_HIP(hipStreamSynchronize(st));
  }
  F3D_HIP(hipEventSynchronize(h->ev1));
  float ms = 0;
  F3D_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (stats) {
    stats->iterations = iters;
    stats->max_dT = maxdt;
    double w = 0;
    for (int d = 0; d < 3; ++d)
      w += 2.0 * (G.n[d] - 1) * (double)G.n[(d + 1) % 3] * G.n[(d + 2) % 3];
    stats->cell_updates = w * iters;
    stats->sweep_ms = ms;
    stats->updates_computed = stats->cell_updates;
    if (done_ctr) {
      unsigned long long done[256];
      F3D_HIP(hipMemcpy(done, done_ctr, sizeof done, hipMemcpyDeviceToHost));
      double sum = 0;
      for (int k = 0; k < 256; ++k) sum += (double)done[k];
      // the counters hold what the locked iterations computed; the iterations
      // before the lock start computed everything
      const int unlocked = iters < locked_from ? iters : locked_from;
      stats->updates_computed = w * unlocked + sum;
    }
    const bool ellip = model->kind == FOLD3D_MODEL_ELLIP;
    const int params = ellip ? (model->speed_field ? 1 : 0) + (model->theta ? 4 : 0) : 0;
    stats->bytes_moved = stats->updates_computed * (2 + params) *
                         (o.precision == FOLD3D_F64 ? 8.0 : 4.0);
    stats->nodes_held = 0;
  }
  if (o.tol > 0 && !(maxdt < o.tol)) {
    set_error("fold3d_solve: max|dT| = %g >= tol %g after %d iterations", maxdt, o.tol, iters);
    return FOLD3D_E_NOT_CONVERGED;
  }
  return FOLD3D_OK;
}

static int check_batch(fold3d_handle *h, int count, const fold3d_instance *inst,
                       const fold3d_opts &o) {
  F3D_CHECK(h && inst, FOLD3D_E_INVALID, "fold3d_solve_batched: NULL handle or instances");
  F3D_CHECK(count >= 1, FOLD3D_E_INVALID, "fold3d_solve_batched: count must be >= 1");
  F3D_CHECK(o.memory == FOLD3D_DEVICE, FOLD3D_E_INVALID,
            "fold3d_solve_batched: instances must be device-resident");
  F3D_CHECK(o.precision == FOLD3D_F64 || o.precision == FOLD3D_F32, FOLD3D_E_INVALID,
            "fold3d_solve_batched: bad precision %d", o.precision);
  F3D_CHECK(o.max_iters >= 1 && o.tol >= 0 && std::isfinite(o.tol), FOLD3D_E_INVALID,
            "fold3d_solve_batched: bad max_iters or tol");
  for (int i = 0; i < count; ++i) {
    const fold3d_instance &x = inst[i];
    F3D_CHECK(x.T != nullptr, FOLD3D_E_INVALID, "fold3d_solve_batched: instance %d has no T", i);
    F3D_CHECK(x.n[0] >= 1 && x.n[1] >= 1 && x.n[2] >= 1, FOLD3D_E_INVALID,
              "fold3d_solve_batched: instance %d grid too small", i);
    F3D_CHECK(grid_bytes_fit(x.n[0], x.n[1], x.n[2]), FOLD3D_E_NOMEM,
              "fold3d_solve_batched: instance %d does not fit in device memory", i);
    F3D_CHECK(x.h[0] > 0 && x.h[1] > 0 && x.h[2] > 0 && std::isfinite(x.h[0]) &&
                  std::isfinite(x.h[1]) && std::isfinite(x.h[2]),
              FOLD3D_E_INVALID, "fold3d_solve_batched: instance %d spacings must be positive", i);
    const int nf = (x.theta != nullptr) + (x.phi != nullptr) + (x.ra != nullptr) + (x.rb != nullptr);
    F3D_CHECK(nf == 0 || nf == 4, FOLD3D_E_INVALID,
              "fold3d_solve_batched: instance %d: give all of theta, phi, ra, rb or none", i);
    F3D_CHECK(x.speed_field != nullptr || (x.speed > 0 && std::isfinite(x.speed)),
              FOLD3D_E_INVALID, "fold3d_solve_batched: instance %d speed must be positive", i);
    if (nf == 0) {
      const double *M = x.M;
      bool fin = true;
      for (int k = 0; k < 6; ++k) fin = fin && std::isfinite(M[k]);
      const double m2 = M[0] * M[1] - M[3] * M[3];
      const double m3 = M[0] * (M[1] * M[2] - M[5] * M[5]) - M[3] * (M[3] * M[2] - M[5] * M[4]) +
                        M[4] * (M[3] * M[5] - M[1] * M[4]);
      F3D_CHECK(fin && M[0] > 0 && m2 > 0 && m3 > 0, FOLD3D_E_INVALID,
                "fold3d_solve_batched: instance %d metric is not finite positive definite", i);
    }
    // host-side checks first; then the pointers (a runtime query)
    F3D_CHECK(is_device_ptr(x.T) && (!x.speed_field || is_device_ptr(x.speed_field)) &&
                  (!x.theta || (is_device_ptr(x.theta) && is_device_ptr(x.phi) &&
                                is_device_ptr(x.ra) && is_device_ptr(x.rb))),
              FOLD3D_E_INVALID, "fold3d_solve_batched: instance %d: fields must be device memory",
              i);
  }
  return FOLD3D_OK;
}

// The instances cut into `k` bands of about equal work (nodes), largest
// instances first; band b gets the instances of positions [cut[b], cut[b+1])
// of `order`.
static void batch_bands(int count, const fold3d_instance *inst, int k, std::vector<int> &order,
                        std::vector<int> &cut) {
  order.resize(count);
  std::iota(order.begin(), order.end(), 0);
  auto nodes = [&](int i) { return (double)inst[i].n[0] * inst[i].n[1] * inst[i].n[2]; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nodes(a) > nodes(b); });
  double total = 0;
  for (int i = 0; i < count; ++i) total += nodes(i);
  cut.assign(1, 0);
  double acc = 0;
  for (int p = 0; p < count && (int)cut.size() < k; ++p) {
    acc += nodes(order[p]);
    if (acc >= total * cut.size() / k && p + 1 < count) cut.push_back(p + 1);
  }
  cut.push_back(count);
}

static int run_batch(fold3d_handle *h, int count, const fold3d_instance *inst,
                     const fold3d_opts &o, int groups, int blocks_per_cu, int rows_per_pull,
                     int *iterations_out, double *max_dt_out, double *ms_out) {
  int cus = 0;
  F3D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const int use_graph = o.engine != FOLD3D_ENGINE_PLANE;
  const int blocks = cus * blocks_per_cu;
  {
    const int rc = ensure_scratch(h, batch_scratch_bytes(count, inst, o.precision));
    if (rc != FOLD3D_OK) return rc;
  }
  hipStream_t st = (hipStream_t)o.stream;
  auto solve = [&](int n, const fold3d_instance *in, void *scr, int *its, double *mds,
                   double *ms) {
    return o.precision == FOLD3D_F64
               ? batch_solve<double>(st, n, in, o.max_iters, o.tol, blocks, groups,
                                     rows_per_pull, use_graph, scr, its, mds, ms)
               : batch_solve<float>(st, n, in, o.max_iters, o.tol, blocks, groups,
                                    rows_per_pull, use_graph, scr, its, mds, ms);
  };
  const int k = std::min(h->batch_streams, count);
  if (k <= 1) return solve(count, inst, h->d_scratch, iterations_out, max_dt_out, ms_out);

  // k sub-batches at once, each on a stream of its own from a host thread of
  // its own: while one waits on its host-side convergence check or runs a
  // plane step that few of its instances have, the others fill the GPU.
  // Every instance's values are computed exactly as in one batch.
  std::vector<int> order, cut;
  batch_bands(count, inst, k, order, cut);
  const int nb = (int)cut.size() - 1;
  std::vector<std::vector<fold3d_instance>> sub(nb);
  std::vector<std::vector<int>> its(nb);
  std::vector<std::vector<double>> mds(nb);
  std::vector<char *> scr(nb);
  std::vector<int> rcs(nb, FOLD3D_OK);
  std::vector<std::string> errs(nb);
  char *p = (char *)h->d_scratch;
  for (int b = 0; b < nb; ++b) {
    for (int q = cut[b]; q < cut[b + 1]; ++q) sub[b].push_back(inst[order[q]]);
    its[b].assign(sub[b].size(), 0);
    mds[b].assign(sub[b].size(), -1.0);
    scr[b] = p;
    p += batch_scratch_bytes((int)sub[b].size(), sub[b].data(), o.precision);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms_out) {
    F3D_HIP(hipEventCreate(&e0));
    F3D_HIP(hipEventCreate(&e1));
    F3D_HIP(hipEventRecord(e0, st));
  }
  auto work = [&](int b) {
    if (hipSetDevice(h->device) != hipSuccess) {
      rcs[b] = FOLD3D_E_HIP;
      errs[b] = "fold3d_solve_batched: hipSetDevice failed in a sub-batch thread";
      return;
    }
    rcs[b] = solve((int)sub[b].size(), sub[b].data(), scr[b], its[b].data(), mds[b].data(),
                   nullptr);
    if (rcs[b] != FOLD3D_OK) errs[b] = fold3d_last_error();
  };
  std::vector<std::thread> pool;
  for (int b = 1; b < nb; ++b) pool.emplace_back(work, b);
  work(0);
  for (auto &t : pool) t.join();
  int rc = FOLD3D_OK;
  for (int b = 0; b < nb; ++b) {
    for (int q = cut[b]; q < cut[b + 1]; ++q) {
      if (iterations_out) iterations_out[order[q]] = its[b][q - cut[b]];
      if (max_dt_out) max_dt_out[order[q]] = mds[b][q - cut[b]];
    }
    // the first failure wins, and a real error over a non-convergence
    if (rcs[b] != FOLD3D_OK &&
        (rc == FOLD3D_OK || (rc == FOLD3D_E_NOT_CONVERGED && rcs[b] != rc))) {
      rc = rcs[b];
      set_error("%s", errs[b].c_str());
    }
  }
  if (ms_out) {
    float ms = 0;
    if (hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
      *ms_out = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return rc;
}

int fold3d_solve_batched(fold3d_handle *h, int count, const fold3d_instance *inst,
                         const fold3d_opts *opts, int *iterations_out, double *max_dt_out,
                         double *ms_out) {
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  const int rc = check_batch(h, count, inst, o);
  if (rc != FOLD3D_OK) return rc;
  F3D_HIP(hipSetDevice(h->device));
  return run_batch(h, count, inst, o, h->batch_groups, h->batch_blocks_per_cu,
                   h->batch_rows_per_pull, iterations_out, max_dt_out, ms_out);
}

int fold3d_autotune_batched(fold3d_handle *h, int count, const fold3d_instance *inst,
                            const fold3d_opts *opts, int iterations, float *ms_out) {
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  int rc = check_batch(h, count, inst, o);
  if (rc != FOLD3D_OK) return rc;
  F3D_CHECK(iterations >= 1, FOLD3D_E_INVALID, "fold3d_autotune_batched: iterations must be >= 1");
  F3D_HIP(hipSetDevice(h->device));
  // scratch copies of every T (the caller's fields are not modified)
  const size_t r = o.precision == FOLD3D_F64 ? 8 : 4;
  std::vector<size_t> off(count + 1, 0);
  for (int i = 0; i < count; ++i) {
    const Grid G(inst[i].n[0], inst[i].n[1], inst[i].n[2]);
    off[i + 1] = (off[i] + (size_t)G.count * r + 255) & ~(size_t)255;
  }
  char *work = nullptr;
  F3D_HIP(hipMalloc((void **)&work, off[count]));
  std::vector<fold3d_instance> copy(inst, inst + count);
  for (int i = 0; i < count; ++i) copy[i].T = work + off[i];
  hipStream_t st = (hipStream_t)o.stream;
  o.max_iters = iterations;
  o.tol = 0.0;
  // (groups, blocks per CU, rows per pull): static schedules and queue pulls
  static const int kConfigs[][3] = {{1, 16, -1}, {1, 32, -1}, {1, 64, -1}, {2, 16, -1},
                                    {2, 32, -1}, {2, 64, -1}, {4, 16, -1}, {1, 8, 8},
                                    {2, 8, 8}};
  float best_ms = 0;
  int best = -1;
  for (int pass = 0; pass < 2 && rc == FOLD3D_OK; ++pass) {  // the first pass warms up
    for (int c = 0; c < (int)(sizeof kConfigs / sizeof kConfigs[0]); ++c) {
      for (int i = 0; i < count && rc == FOLD3D_OK; ++i) {
        const Grid G(inst[i].n[0], inst[i].n[1], inst[i].n[2]);
        if (hipMemcpyAsync(copy[i].T, inst[i].T, (size_t)G.count * r, hipMemcpyDeviceToDevice,
                           st) != hipSuccess)
          rc = FOLD3D_E_HIP;
      }
      double ms = 0;
      if (rc == FOLD3D_OK)
        rc = run_batch(h, count, copy.data(), o, kConfigs[c][0], kConfigs[c][1], kConfigs[c][2],
                       nullptr, nullptr, &ms);
      if (rc != FOLD3D_OK) break;
      if (pass == 1 && (best < 0 || ms < best_ms)) {
        best = c;
        best_ms = (float)ms;
      }
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(work);
  if (rc != FOLD3D_OK) return rc;
  h->batch_groups = kConfigs[best][0];
  h->batch_blocks_per_cu = kConfigs[best][1];
  h->batch_rows_per_pull = kConfigs[best][2];
  if (ms_out) *ms_out = best_ms;
  return FOLD3D_OK;
}

}  // extern "C"

// Checks shared by the adjoint entry points: model, grid, options.
static int check_adjoint_common(const char *who, int nx, int ny, int nz, double dx, double dy,
                                double dz, const fold3d_model *model, void *grad_slowness,
                                void *grad_theta, const fold3d_opts &o) {
  F3D_CHECK(model && grad_slowness, FOLD3D_E_INVALID, "%s: NULL model or grad_slowness", who);
  F3D_CHECK(model->kind == FOLD3D_MODEL_ELLIP, FOLD3D_E_INVALID,
            "%s: the adjoint is for the elliptic model", who);
  F3D_CHECK(nx >= 1 && ny >= 1 && nz >= 1, FOLD3D_E_INVALID, "%s: grid too small", who);
  F3D_CHECK(dx > 0 && dy > 0 && dz > 0 && std::isfinite(dx) && std::isfinite(dy) &&
                std::isfinite(dz),
            FOLD3D_E_INVALID, "%s: bad spacings", who);
  F3D_CHECK(grad_theta == nullptr || model->theta != nullptr, FOLD3D_E_INVALID,
            "%s: a tilt gradient needs per-node tilt fields", who);
  F3D_CHECK(model->speed_field != nullptr || (model->speed > 0 && std::isfinite(model->speed)),
            FOLD3D_E_INVALID, "%s: speed must be positive", who);
  const int nf = (model->theta != nullptr) + (model->phi != nullptr) + (model->ra != nullptr) +
                 (model->rb != nullptr);
  F3D_CHECK(nf == 0 || nf == 4, FOLD3D_E_INVALID,
            "%s: give all four of theta, phi, ra, rb or none", who);
  if (nf == 0) {
    const double *M = model->M;
    bool fin = true;
    for (int k = 0; k < 6; ++k) fin = fin && std::isfinite(M[k]);
    const double m2 = M[0] * M[1] - M[3] * M[3];
    const double m3 = M[0] * (M[1] * M[2] - M[5] * M[5]) - M[3] * (M[3] * M[2] - M[5] * M[4]) +
                      M[4] * (M[3] * M[5] - M[1] * M[4]);
    F3D_CHECK(fin && M[0] > 0 && m2 > 0 && m3 > 0, FOLD3D_E_INVALID,
              "%s: metric M is not symmetric positive definite", who);
  }
  F3D_CHECK(o.memory == FOLD3D_DEVICE, FOLD3D_E_INVALID, "%s: T must be on the device", who);
  F3D_CHECK(o.precision == FOLD3D_F64 || o.precision == FOLD3D_F32, FOLD3D_E_INVALID,
            "%s: bad precision %d", who, o.precision);
  F3D_CHECK(o.max_iters >= 1 && o.tol >= 0, FOLD3D_E_INVALID, "%s: bad max_iters/tol", who);
  return FOLD3D_OK;
}

// Device-side checks of one source: device arrays, receivers at interior nodes
// (their residuals are scattered by index).
static int check_adjoint_source(const char *who, int k, const Grid &G, const void *T, int nrec,
                                const int *rec_idx, const void *rec_res, double tol,
                                hipStream_t st) {
  F3D_CHECK(T != nullptr, FOLD3D_E_INVALID, "%s: source %d: NULL T", who, k);
  F3D_CHECK(nrec >= 0 && (nrec == 0 || (rec_idx && rec_res)), FOLD3D_E_INVALID,
            "%s: source %d: bad receiver arrays", who, k);
  F3D_CHECK(tol >= 0, FOLD3D_E_INVALID, "%s: source %d: negative tolerance", who, k);
  F3D_CHECK_DEV(T, "adjoint: T");
  if (nrec == 0) return FOLD3D_OK;
  F3D_CHECK_DEV(rec_idx, "adjoint: rec_idx");
  F3D_CHECK_DEV(rec_res, "adjoint: rec_res");
  std::vector<int> idx(nrec);
  F3D_HIP(hipMemcpyAsync(idx.data(), rec_idx, sizeof(int) * nrec, hipMemcpyDeviceToHost, st));
  F3D_HIP(hipStreamSynchronize(st));
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
  for (int r = 0; r < nrec; ++r) {
    const long long x = idx[r];
    const long long i = x % G.stride[1], j = (x / G.stride[1]) % (ny + 2), kk = x / G.stride[2];
    F3D_CHECK(x >= 0 && x < G.count && i >= 1 && i <= nx && j >= 1 && j <= ny && kk >= 1 &&
                  kk <= nz,
              FOLD3D_E_INVALID, "%s: source %d: receiver %d index %lld is not an interior node",
              who, k, r, x);
  }
  return FOLD3D_OK;
}

// Runs the checked adjoint of `nsrc` sources; per-source outputs.
static int run_adjoint(fold3d_handle *h, const char *who, int nsrc, const AdjointSource *srcs,
                       int nx, int ny, int nz, double dx, double dy, double dz,
                       const fold3d_model *model, const fold3d_opts &o, void *grad_slowness,
                       void *grad_theta, int *iters, double *maxd, long long *held,
                       double *ms_out) {
  const Grid G(nx, ny, nz);
  hipStream_t st = (hipStream_t)o.stream;
  const int rc0 = ensure_scratch(h, adjoint_scratch_bytes(G, o.precision, nsrc));
  if (rc0 != FOLD3D_OK) return rc0;
  const double hsp[3] = {dx, dy, dz};
  F3D_HIP(hipEventRecord(h->ev0, st));
  const int rc =
      o.precision == FOLD3D_F64
          ? adjoint_solve<double>(st, nsrc, srcs, G, *model, hsp, o.max_iters,
                                  (double *)grad_slowness, (double *)grad_theta, h->d_scratch,
                                  iters, maxd, held)
          : adjoint_solve<float>(st, nsrc, srcs, G, *model, hsp, o.max_iters,
                                 (float *)grad_slowness, (float *)grad_theta, h->d_scratch, iters,
                                 maxd, held);
  F3D_HIP(hipEventRecord(h->ev1, st));
  if (rc != FOLD3D_OK) return rc;
  F3D_HIP(hipEventSynchronize(h->ev1));
  float ms = 0;
  F3D_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (ms_out) *ms_out = ms;
  for (int k = 0; k < nsrc; ++k)
    if (!(maxd[k] <= srcs[k].tol)) {
      set_error("%s: source %d: adjoint change %g > tol %g after %d sweeps", who, k, maxd[k],
                srcs[k].tol, iters[k]);
      return FOLD3D_E_NOT_CONVERGED;
    }
  return FOLD3D_OK;
}

extern "C" {

int fold3d_adjoint(fold3d_handle *h, const void *T, int nx, int ny, int nz, double dx,
                   double dy, double dz, const fold3d_model *model, int nrec,
                   const int *rec_idx, const void *rec_res, const fold3d_opts *opts,
                   void *grad_slowness, void *grad_theta, fold3d_stats *stats) {
  static const char *who = "fold3d_adjoint";
  F3D_CHECK(h && T, FOLD3D_E_INVALID, "fold3d_adjoint: NULL handle or T");
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  int rc = check_adjoint_common(who, nx, ny, nz, dx, dy, dz, model, grad_slowness, grad_theta, o);
  if (rc != FOLD3D_OK) return rc;
  F3D_CHECK(nrec >= 0 && (nrec == 0 || (rec_idx && rec_res)), FOLD3D_E_INVALID,
            "fold3d_adjoint: bad receiver arrays");
  F3D_HIP(hipSetDevice(h->device));
  F3D_CHECK_DEV(grad_slowness, "fold3d_adjoint: grad_slowness");
  F3D_CHECK_DEV(grad_theta, "fold3d_adjoint: grad_theta");
  if (check_model_fields(model) != FOLD3D_OK) return FOLD3D_E_INVALID;
  const Grid G(nx, ny, nz);
  rc = check_adjoint_source(who, 0, G, T, nrec, rec_idx, rec_res, o.tol, (hipStream_t)o.stream);
  if (rc != FOLD3D_OK) return rc;
  const AdjointSource src{T, nrec, rec_idx, rec_res, o.tol};
  int iters = 0;
  double maxd = -1, ms = 0;
  long long held = 0;
  rc = run_adjoint(h, who, 1, &src, nx, ny, nz, dx, dy, dz, model, o, grad_slowness, grad_theta,
                   &iters, &maxd, &held, &ms);
  if (stats && (rc == FOLD3D_OK || rc == FOLD3D_E_NOT_CONVERGED)) {
    stats->iterations = iters;
    stats->max_dT = maxd;
    stats->cell_updates = 0;
    stats->sweep_ms = ms;
    stats->updates_computed = 0;
    stats->bytes_moved = 0;
    stats->nodes_held = (double)held;
  }
  return rc;
}

int fold3d_adjoint_batched(fold3d_handle *h, int count, const fold3d_adjoint_source *src,
                           int nx, int ny, int nz, double dx, double dy, double dz,
                           const fold3d_model *model, const fold3d_opts *opts,
                           void *grad_slowness, void *grad_theta, int *iterations_out,
                           double *max_d_out, long long *held_out, double *ms_out) {
  static const char *who = "fold3d_adjoint_batched";
  F3D_CHECK(h && src, FOLD3D_E_INVALID, "fold3d_adjoint_batched: NULL handle or sources");
  F3D_CHECK(count >= 1 && count <= 4096, FOLD3D_E_INVALID,
            "fold3d_adjoint_batched: count must be 1..4096");
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  int rc = check_adjoint_common(who, nx, ny, nz, dx, dy, dz, model, grad_slowness, grad_theta, o);
  if (rc != FOLD3D_OK) return rc;
  for (int k = 0; k < count; ++k)
    F3D_CHECK(src[k].T && src[k].nrec >= 0 && (src[k].nrec == 0 || (src[k].rec_idx && src[k].rec_res)) &&
                  src[k].tol >= 0,
              FOLD3D_E_INVALID, "fold3d_adjoint_batched: source %d: bad T, receivers or tol", k);
  F3D_HIP(hipSetDevice(h->device));
  F3D_CHECK_DEV(grad_slowness, "fold3d_adjoint_batched: grad_slowness");
  F3D_CHECK_DEV(grad_theta, "fold3d_adjoint_batched: grad_theta");
  if (check_model_fields(model) != FOLD3D_OK) return FOLD3D_E_INVALID;
  const Grid G(nx, ny, nz);
  std::vector<AdjointSource> srcs(count);
  for (int k = 0; k < count; ++k) {
    rc = check_adjoint_source(who, k, G, src[k].T, src[k].nrec, src[k].rec_idx, src[k].rec_res,
                              src[k].tol, (hipStream_t)o.stream);
    if (rc != FOLD3D_OK) return rc;
    srcs[k] = AdjointSource{src[k].T, src[k].nrec, src[k].rec_idx, src[k].rec_res, src[k].tol};
  }
  std::vector<int> iters(count, 0);
  std::vector<double> maxd(count, -1.0);
  std::vector<long long> held(count, 0);
  double ms = 0;
  rc = run_adjoint(h, who, count, srcs.data(), nx, ny, nz, dx, dy, dz, model, o, grad_slowness,
                   grad_theta, iters.data(), maxd.data(), held.data(), &ms);
  if (rc == FOLD3D_OK || rc == FOLD3D_E_NOT_CONVERGED) {
    for (int k = 0; k < count; ++k) {
      if (iterations_out) iterations_out[k] = iters[k];
      if (max_d_out) max_d_out[k] = maxd[k];
      if (held_out) held_out[k] = held[k];
    }
    if (ms_out) *ms_out = ms;
  }
  return rc;
}

int fold3d_set_tuning(fold3d_handle *h, int model_kind, int groups) {
  F3D_CHECK(h != nullptr, FOLD3D_E_INVALID, "fold3d_set_tuning: NULL handle");
  if (model_kind == FOLD3D_MODEL_FOLD) {
    F3D_CHECK(groups == 0 || groups == 2 || groups == 4 || groups == 8, FOLD3D_E_INVALID,
              "fold3d_set_tuning: fold-model groups must be 0, 2, 4 or 8");
    h->split_groups = groups == 0 ? 4 : groups;
  } else {
    F3D_CHECK(model_kind == FOLD3D_MODEL_ELLIP, FOLD3D_E_INVALID,
              "fold3d_set_tuning: unknown model kind %d", model_kind);
    F3D_CHECK(groups == 0 || groups == 1 || groups == 2 || groups == 4, FOLD3D_E_INVALID,
              "fold3d_set_tuning: elliptic-model groups must be 0, 1, 2 or 4");
    h->ellip_groups = groups == 0 ? 2 : groups;
  }
  return FOLD3D_OK;
}

int fold3d_set_lock_start(fold3d_handle *h, int iteration) {
  F3D_CHECK(h != nullptr, FOLD3D_E_INVALID, "fold3d_set_lock_start: NULL handle");
  F3D_CHECK(iteration >= -1, FOLD3D_E_INVALID,
            "fold3d_set_lock_start: iteration must be >= 0 (or -1 for the default)");
  h->lock_from = iteration < 0 ? kDefaultLockFrom : iteration;
  return FOLD3D_OK;
}

int fold3d_get_lock_start(fold3d_handle *h, int *iteration) {
  F3D_CHECK(h && iteration, FOLD3D_E_INVALID, "fold3d_get_lock_start: NULL argument");
  *iteration = h->lock_from;
  return FOLD3D_OK;
}

int fold3d_get_tuning(fold3d_handle *h, int model_kind, int *groups) {
  F3D_CHECK(h && groups, FOLD3D_E_INVALID, "fold3d_get_tuning: NULL argument");
  *groups = model_kind == FOLD3D_MODEL_FOLD ? h->split_groups : h->ellip_groups;
  return FOLD3D_OK;
}

int fold3d_initialiser_device(fold3d_handle *h, double *T, int nx, int ny, int nz, double dx,
                              double dy, double dz, void *stream, long long *n_exact) {
  F3D_CHECK(h && T, FOLD3D_E_INVALID, "fold3d_initialiser_device: NULL handle or T");
  F3D_CHECK(nx >= 1 && ny >= 1 && nz >= 1, FOLD3D_E_INVALID,
            "fold3d_initialiser_device: bad grid %d x %d x %d", nx, ny, nz);
  F3D_CHECK(dx > 0 && dy > 0 && dz > 0 && std::isfinite(dx) && std::isfinite(dy) &&
                std::isfinite(dz),
            FOLD3D_E_INVALID, "fold3d_initialiser_device: spacings must be positive");
  F3D_HIP(hipSetDevice(h->device));
  F3D_CHECK_DEV(T, "fold3d_initialiser_device: T");
  fold3d_set_grid(nx, ny, nz);
  return device_initialiser((hipStream_t)stream, T, nx, ny, nz, dx, dy, dz, n_exact);
}

int fold3d_point_source(fold3d_handle *h, void *T, int nx, int ny, int nz, double dx,
                        double dy, double dz, const double src[3], const double M[6],
                        double speed, double radius, int precision, void *stream) {
  F3D_CHECK(h && T && src && M, FOLD3D_E_INVALID, "fold3d_point_source: NULL argument");
  F3D_CHECK(nx >= 1 && ny >= 1 && nz >= 1, FOLD3D_E_INVALID, "fold3d_point_source: bad grid");
  F3D_CHECK(dx > 0 && dy > 0 && dz > 0 && speed > 0 && radius >= 0 && std::isfinite(speed),
            FOLD3D_E_INVALID, "fold3d_point_source: spacings, speed and radius must be positive");
  F3D_CHECK(precision == FOLD3D_F64 || precision == FOLD3D_F32, FOLD3D_E_INVALID,
            "fold3d_point_source: bad precision %d", precision);
  const double m2 = M[0] * M[1] - M[3] * M[3];
  const double m3 = M[0] * (M[1] * M[2] - M[5] * M[5]) - M[3] * (M[3] * M[2] - M[5] * M[4]) +
                    M[4] * (M[3] * M[5] - M[1] * M[4]);
  F3D_CHECK(M[0] > 0 && m2 > 0 && m3 > 0, FOLD3D_E_INVALID,
            "fold3d_point_source: metric is not positive definite");
  F3D_HIP(hipSetDevice(h->device));
  F3D_CHECK_DEV(T, "fold3d_point_source: T");
  double N[6];
  sym3_inverse(M, N);
  const double hh[3] = {dx, dy, dz};
  return point_source_start((hipStream_t)stream, T, precision, nx, ny, nz, hh, src, N, speed,
                            radius);
}

int fold3d_set_batch_tuning(fold3d_handle *h, int groups, int blocks_per_cu, int rows_per_pull) {
  F3D_CHECK(h != nullptr, FOLD3D_E_INVALID, "fold3d_set_batch_tuning: NULL handle");
  F3D_CHECK(groups == 0 || groups == 1 || groups == 2 || groups == 4, FOLD3D_E_INVALID,
            "fold3d_set_batch_tuning: groups must be 0, 1, 2 or 4");
  F3D_CHECK(blocks_per_cu >= 0 && blocks_per_cu <= 64, FOLD3D_E_INVALID,
            "fold3d_set_batch_tuning: blocks_per_cu must be 0..64");
  F3D_CHECK(rows_per_pull >= -1 && rows_per_pull <= 64, FOLD3D_E_INVALID,
            "fold3d_set_batch_tuning: rows_per_pull must be -1..64");
  h->batch_groups = groups == 0 ? 1 : groups;
  h->batch_blocks_per_cu = blocks_per_cu == 0 ? 64 : blocks_per_cu;
  h->batch_rows_per_pull = rows_per_pull;
  return FOLD3D_OK;
}

int fold3d_set_batch_streams(fold3d_handle *h, int streams) {
  F3D_CHECK(h != nullptr, FOLD3D_E_INVALID, "fold3d_set_batch_streams: NULL handle");
  F3D_CHECK(streams >= 0 && streams <= kMaxBatchStreams, FOLD3D_E_INVALID,
            "fold3d_set_batch_streams: streams must be 0..%d", kMaxBatchStreams);
  h->batch_streams = streams == 0 ? kDefaultBatchStreams : streams;
  return FOLD3D_OK;
}

int fold3d_get_batch_streams(fold3d_handle *h, int *streams) {
  F3D_CHECK(h && streams, FOLD3D_E_INVALID, "fold3d_get_batch_streams: NULL argument");
  *streams = h->batch_streams;
  return FOLD3D_OK;
}

int fold3d_get_batch_tuning(fold3d_handle *h, int *groups, int *blocks_per_cu,
                            int *rows_per_pull) {
  F3D_CHECK(h && groups && blocks_per_cu && rows_per_pull, FOLD3D_E_INVALID,
            "fold3d_get_batch_tuning: NULL argument");
  *groups = h->batch_groups;
  *blocks_per_cu = h->batch_blocks_per_cu;
  *rows_per_pull = h->batch_rows_per_pull;
  return FOLD3D_OK;
}

int fold3d_autotune(fold3d_handle *h, const void *T, int nx, int ny, int nz, double dx,
                    double dy, double dz, const fold3d_model *model, const fold3d_opts *opts,
                    int iterations, float *ms_out) {
  F3D_CHECK(h && T && model, FOLD3D_E_INVALID, "fold3d_autotune: NULL handle, T or model");
  F3D_CHECK(iterations >= 1, FOLD3D_E_INVALID, "fold3d_autotune: iterations must be >= 1");
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  F3D_CHECK(o.memory == FOLD3D_DEVICE, FOLD3D_E_INVALID, "fold3d_autotune: T must be on the device");
  F3D_CHECK(nx >= 1 && ny >= 1 && nz >= 1, FOLD3D_E_INVALID, "fold3d_autotune: grid too small");
  const Grid G(nx, ny, nz);
  const size_t bytes = (size_t)G.count * (o.precision == FOLD3D_F64 ? 8 : 4);
  F3D_HIP(hipSetDevice(h->device));
  F3D_CHECK_DEV(T, "fold3d_autotune: T");
  void *work = nullptr;
  F3D_HIP(hipMalloc(&work, bytes));
  hipStream_t st = (hipStream_t)o.stream;
  o.max_iters = iterations;
  o.tol = 0.0;
  o.engine = FOLD3D_ENGINE_SPLIT;
  const bool fold = model->kind == FOLD3D_MODEL_FOLD;
  static const int kFoldCandidates[3] = {2, 4, 8};
  static const int kEllipCandidates[3] = {1, 2, 4};
  const int *cand = fold ? kFoldCandidates : kEllipCandidates;
  int &knob = fold ? h->split_groups : h->ellip_groups;
  int best = knob;
  float best_ms = -1.0f;
  int rc = FOLD3D_OK;
  for (int r = 0; r < 2 && rc == FOLD3D_OK; ++r) {  // the first pass warms up
    for (int c = 0; c < 3 && rc == FOLD3D_OK; ++c) {
      knob = cand[c];
      if (hipMemcpyAsync(work, T, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) {
        rc = FOLD3D_E_HIP;
        break;
      }
      fold3d_stats s;
      rc = fold3d_solve(h, work, nx, ny, nz, dx, dy, dz, model, &o, &s);
      if (rc == FOLD3D_OK && r == 1 && (best_ms < 0 || s.sweep_ms < best_ms)) {
        best_ms = (float)s.sweep_ms;
        best = cand[c];
      }
    }
  }
  knob = best;
  (void)hipStreamSynchronize(st);
  (void)hipFree(work);
  if (ms_out) *ms_out = best_ms;
  return rc;
}

int fold3d_plane_step(fold3d_handle *h, int direction, double *T, int nx, int ny,
                      int nz, int plane, double dzz, double dxx, double dyy,
                      double dxy, double dxz, double dyz, double F, double ay,
                      double ax, double az, void *stream) {
  F3D_CHECK(h && T, FOLD3D_E_INVALID, "fold3d_plane_step: NULL handle or T");
  F3D_CHECK(direction >= 0 && direction < 6, FOLD3D_E_INVALID,
            "fold3d_plane_step: direction %d not in [0,6)", direction);
  F3D_CHECK(nx >= 3 && ny >= 3 && nz >= 3, FOLD3D_E_INVALID,
            "fold3d_plane_step: grid too small");
  const Grid G(nx, ny, nz);
  const Frame &fr = kFrames[direction];
  const int n = G.n[fr.m];
  const bool ok = fr.s > 0 ? (plane >= 1 && plane <= n - 1) : (plane >= 2 && plane <= n);
  F3D_CHECK(ok, FOLD3D_E_INVALID, "fold3d_plane_step: plane %d out of range for direction %d",
            plane, direction);
  F3D_HIP(hipSetDevice(h->device));
  F3D_CHECK_DEV(T, "fold3d_plane_step: T");
  FoldFrame<double> f{F, ax, ay, az, dzz, dxx, dyy, dxy, dxz, dyz};
  return fold_plane_step_f64((hipStream_t)stream, T, G, direction, plane, f);
}

#define F3D_PLANE_STEP(name, dir)                                                  \
  int name(fold3d_handle *h, double *T, int nx, int ny, int nz, int plane, double dzz,  \
           double dxx, double dyy, double dxy, double dxz, double dyz, double F,         \
           double ay, double ax, double az, void *stream) {                             \
    return fold3d_plane_step(h, dir, T, nx, ny, nz, plane, dzz, dxx, dyy, dxy, dxz, dyz, \
                             F, ay, ax, az, stream);                                    \
  }
F3D_PLANE_STEP(fold3d_function_x_down, FOLD3D_X_DOWN)
F3D_PLANE_STEP(fold3d_function_x_up, FOLD3D_X_UP)
F3D_PLANE_STEP(fold3d_function_k_up, FOLD3D_K_UP)
F3D_PLANE_STEP(fold3d_function_k_down, FOLD3D_K_DOWN)
F3D_PLANE_STEP(fold3d_function_y_up, FOLD3D_Y_UP)
F3D_PLANE_STEP(fold3d_function_y_down, FOLD3D_Y_DOWN)

int fold3d_device_alloc(void **ptr, size_t bytes) {
  F3D_CHECK(ptr != nullptr, FOLD3D_E_INVALID, "fold3d_device_alloc: ptr is NULL");
  F3D_HIP(hipMalloc(ptr, bytes));
  return FOLD3D_OK;
}
int fold3d_device_free(void *ptr) {
  F3D_HIP(hipFree(ptr));
  return FOLD3D_OK;
}
int fold3d_copy_to_device(void *dst, const void *src, size_t bytes) {
  F3D_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FOLD3D_OK;
}
int fold3d_copy_to_host(void *dst, const void *src, size_t bytes) {
  F3D_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FOLD3D_OK;
}
int fold3d_synchronize(void *stream) {
  F3D_HIP(hipStreamSynchronize((hipStream_t)stream));
  return FOLD3D_OK;
}

const char *fold3d_last_error(void);

// Reference-signature plane steps (AN_ISO_1_OpenACC.c:180-1092): synchronous,
// on the grid of the last ImplicitInitialiser()/fold3d_set_grid() call.  The
// reference's functions return void, so a failure is reported on stderr.
static void dropin_step(int dir, double *T, int plane, double dzz, double dxx,
                        double dyy, double dxy, double dxz, double dyz, double F,
                        double ay, double ax, double az) {
  int nx, ny, nz;
  fold3d_get_grid(&nx, &ny, &nz);
  static fold3d_handle *h = nullptr;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (!h || h->device != dev) {
    if (h) fold3d_destroy(h);
    if (fold3d_create(&h, dev) != FOLD3D_OK) {
      fprintf(stderr, "fold3d: %s\n", fold3d_last_error());
      return;
    }
  }
  int rc = fold3d_plane_step(h, dir, T, nx, ny, nz, plane, dzz, dxx, dyy, dxy, dxz,
                             dyz, F, ay, ax, az, nullptr);
  if (rc == FOLD3D_OK) rc = fold3d_synchronize(nullptr);
  if (rc != FOLD3D_OK) fprintf(stderr, "fold3d: %s\n", fold3d_last_error());
}

#define F3D_DROPIN(name, dir)                                                     \
  void name(double *T, int p, double dzz, double dxx, double dyy, double dxy,     \
            double dxz, double dyz, double F, double ay, double ax, double az) { \
    dropin_step(dir, T, p, dzz, dxx, dyy, dxy, dxz, dyz, F, ay, ax, az);         \
  }
F3D_DROPIN(function_x_down, FOLD3D_X_DOWN)
F3D_DROPIN(function_x_up, FOLD3D_X_UP)
F3D_DROPIN(function_k_up, FOLD3D_K_UP)
F3D_DROPIN(function_k_down, FOLD3D_K_DOWN)
F3D_DROPIN(function_y_up, FOLD3D_Y_UP)
F3D_DROPIN(function_y_down, FOLD3D_Y_DOWN)

}  // extern "C"
