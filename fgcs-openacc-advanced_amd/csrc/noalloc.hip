// noalloc.hip — fold3d_solve_noalloc: the fold-model solve of fold3d_solve on
// a device field, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).  It takes no
// handle and allocates nothing: its scratch is the file-scope device array
// `scratch` below, one per GPU for the life of the process.
//
// What the scratch holds is fixed, whatever the grid: the constant tables of
// the six sweep directions (built on the device by one small kernel from the
// model and spacings, the same expressions and bits as the host build of
// fold3d_solve) and the kMaxSlots words of the tolerance stop's max|dT|
// (cleared by a kernel each iteration).  The sweeps run the split kernel on
// the caller's field in the reference layout for all six directions, so no
// second field is needed.  Same bits as fold3d_solve.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "common.h"
#include "fold_fast.h"
#include "solver.h"

namespace {

constexpr size_t kTableBytes = 6 * sizeof(fold3d::FoldFast<double>);  // FP32 tables are smaller
constexpr size_t kSlotOffset = (kTableBytes + 255) & ~(size_t)255;
constexpr size_t kScratchBytes = 32 * 1024;
static_assert(kSlotOffset + fold3d::kMaxSlots * sizeof(unsigned long long) <= kScratchBytes,
              "the allocation-free solve's scratch must hold six tables and the max|dT| words");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch[kMaxDevices];  // device address of `scratch` on each device, once looked up

}  // namespace

// The allocation-free solve's scratch (the name is a host-side stand-in in
// host code; kernels get its device address from scratch_address()).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace fold3d {

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
static int scratch_address(void **out) {
  int dev = 0;
  F3D_HIP(hipGetDevice(&dev));
  F3D_CHECK(dev >= 0 && dev < kMaxDevices, FOLD3D_E_UNSUPPORTED,
            "fold3d_solve_noalloc: device %d beyond the %d this library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    F3D_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    F3D_CHECK(p != nullptr, FOLD3D_E_HIP, "fold3d_solve_noalloc: no address for the scratch");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return FOLD3D_OK;
}

// The six directions' tables (both signs each), one thread per direction.
template <typename R>
__global__ void noalloc_consts_kernel(FoldFast<R> *__restrict__ out, double F, double a0,
                                      double a1, double a2, double h0, double h1, double h2) {
  const int d = threadIdx.x;
  if (d < 6) {
    const double a[3] = {a0, a1, a2}, h[3] = {h0, h1, h2};
    fold_fast_consts<R>(d, F, a, h, out[d]);
  }
}

// Waves per node row of the split kernel by grid size: what the autotuner
// picks on the reference workload (8 up to 256^3, 4 at 384^3, 2 from 512^3).
static int noalloc_groups(int nx, int ny, int nz) {
  const double side = std::cbrt((double)nx * ny * nz);
  return side >= 448.0 ? 2 : side >= 320.0 ? 4 : 8;
}

template <typename R>
static int noalloc_run(hipStream_t st, R *T, const Grid &G, const fold3d_model &m,
                       const double h[3], const fold3d_opts &o, char *scr, int *iters,
                       double *maxdt) {
  FoldFast<R> *consts = (FoldFast<R> *)scr;
  unsigned long long *slots = (unsigned long long *)(scr + kSlotOffset);
  noalloc_consts_kernel<R><<<1, 64, 0, st>>>(consts, m.F, m.a[0], m.a[1], m.a[2], h[0], h[1],
                                             h[2]);
  F3D_HIP(hipGetLastError());
  return fold_sweeps_inplace_engine<R>(st, T, G, consts, o.max_iters, o.tol,
                                       noalloc_groups(G.n[0], G.n[1], G.n[2]), slots, iters,
                                       maxdt);
}

}  // namespace fold3d

using namespace fold3d;

extern "C" int fold3d_solve_noalloc(void *T, int nx, int ny, int nz, double dx, double dy,
                                    double dz, const fold3d_model *model,
                                    const fold3d_opts *opts, fold3d_stats *stats) {
  F3D_CHECK(T && model, FOLD3D_E_INVALID, "fold3d_solve_noalloc: NULL T or model");
  F3D_CHECK(nx >= 1 && ny >= 1 && nz >= 1, FOLD3D_E_INVALID,
            "fold3d_solve_noalloc: grid %dx%dx%d too small (need >= 1 per axis)", nx, ny, nz);
  F3D_CHECK(8.0 * (nx + 2.0) * (ny + 2.0) * (nz + 2.0) < 0x1p56, FOLD3D_E_NOMEM,
            "fold3d_solve_noalloc: a %dx%dx%d field does not fit in device memory", nx, ny, nz);
  F3D_CHECK(dx > 0 && dy > 0 && dz > 0 && std::isfinite(dx) && std::isfinite(dy) &&
                std::isfinite(dz),
            FOLD3D_E_INVALID, "fold3d_solve_noalloc: spacings must be positive and finite");
  fold3d_opts o;
  fold3d_default_opts(&o);
  if (opts) o = *opts;
  F3D_CHECK(o.max_iters >= 1, FOLD3D_E_INVALID, "fold3d_solve_noalloc: max_iters must be >= 1");
  F3D_CHECK(o.tol >= 0 && std::isfinite(o.tol), FOLD3D_E_INVALID,
            "fold3d_solve_noalloc: bad tol");
  F3D_CHECK(o.precision == FOLD3D_F64 || o.precision == FOLD3D_F32, FOLD3D_E_INVALID,
            "fold3d_solve_noalloc: bad precision %d", o.precision);
  F3D_CHECK(o.memory == FOLD3D_DEVICE, FOLD3D_E_UNSUPPORTED,
            "fold3d_solve_noalloc: T must be device memory (a host T needs a device copy)");
  F3D_CHECK(o.engine == FOLD3D_ENGINE_AUTO || o.engine == FOLD3D_ENGINE_SPLIT,
            FOLD3D_E_UNSUPPORTED, "fold3d_solve_noalloc: runs the split engine only (engine %d)",
            o.engine);
  F3D_CHECK(model->kind == FOLD3D_MODEL_FOLD, FOLD3D_E_UNSUPPORTED,
            "fold3d_solve_noalloc: the fold model only");
  const double *a = model->a;
  F3D_CHECK(std::isfinite(model->F) && model->F > 0, FOLD3D_E_INVALID,
            "fold3d_solve_noalloc: F must be positive");
  F3D_CHECK(std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]) &&
                a[0] * a[0] + a[1] * a[1] + a[2] * a[2] < model->F * model->F,
            FOLD3D_E_INVALID, "fold3d_solve_noalloc: a must be finite with |a| < F");
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, T) != hipSuccess || attr.type == hipMemoryTypeUnregistered) {
    (void)hipGetLastError();
    set_error("fold3d_solve_noalloc: T must be device memory (got a host or unknown pointer)");
    return FOLD3D_E_INVALID;
  }
  void *scr = nullptr;
  const int rs = scratch_address(&scr);
  if (rs != FOLD3D_OK) return rs;
  const Grid G(nx, ny, nz);
  const double hsp[3] = {dx, dy, dz};
  hipStream_t st = (hipStream_t)o.stream;
  int iters = 0;
  double maxdt = -1;
  const int rc =
      o.precision == FOLD3D_F64
          ? noalloc_run<double>(st, (double *)T, G, *model, hsp, o, (char *)scr, &iters, &maxdt)
          : noalloc_run<float>(st, (float *)T, G, *model, hsp, o, (char *)scr, &iters, &maxdt);
  if (rc != FOLD3D_OK) return rc;
  if (stats) {
    double w = 0;
    for (int d = 0; d < 3; ++d)
      w += 2.0 * (G.n[d] - 1) * (double)G.n[(d + 1) % 3] * G.n[(d + 2) % 3];
    stats->iterations = iters;
    stats->max_dT = maxdt;
    stats->cell_updates = w * iters;
    stats->sweep_ms = 0;  // not timed: that would need events
    stats->updates_computed = stats->cell_updates;
    stats->bytes_moved = stats->cell_updates * 2 * (o.precision == FOLD3D_F64 ? 8.0 : 4.0);
    stats->nodes_held = 0;
  }
  if (o.tol > 0 && !(maxdt < o.tol)) {
    set_error("fold3d_solve_noalloc: max|dT| = %g >= tol %g after %d iterations", maxdt, o.tol,
              iters);
    return FOLD3D_E_NOT_CONVERGED;
  }
  return FOLD3D_OK;
}
