// batch.hip — many independent elliptic-model solves of uneven size in one
// stream of launches.
//
// One launch per (sweep direction d, plane step q) serves every instance that
// still has a plane q in direction d; instances are sorted by their plane count
// along d, so those are a prefix of the sorted list.  A launch runs a fixed
// grid of persistent blocks (one wave each) that pull work items — one row of
// 64 target nodes of one instance — from a per-launch device counter until the
// launch's items run out.  The launches of one iteration (plus the counter and
// maximum resets) are captured once in a hipGraph and replayed; between
// replays the host reads each instance's max|dT| and retires converged
// instances through a device flag, so they cost only a flag test afterwards.
// Plane order and node update are those of the single-instance solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.h"
#include "ellip_local.h"
#include "solver.h"

namespace fold3d {

// Device view of one instance (layout Z, the reference layout).
template <typename R>
struct BatchInst {
  R *T;
  const R *V;                      // speed field or null
  const R *th, *ph, *ra, *rb;      // per-node metric or null
  long long stride[3];
  int n[3];
  R h[3];
  R speed;
  R Nc[6];                         // constant inverse metric (no per-node fields)
};

// Per (direction): instances sorted by plane count, prefix sums of their rows.
struct BatchDir {
  const int *order;       // [count]
  const long long *prefix;  // [count + 1], rows of order[0..k)
};

template <typename R>
__global__ __launch_bounds__(64) void batch_plane_kernel(
    const BatchInst<R> *__restrict__ inst, BatchDir bd, int dir, int q, long long total,
    int active, unsigned int *__restrict__ counter, const int *__restrict__ done,
    unsigned long long *__restrict__ max_dt) {
  __shared__ R s[3][64 + 2];
  __shared__ long long item_s;
  const Frame fr = kFrames[dir];
  const int tx = threadIdx.x;
  for (;;) {
    if (tx == 0) item_s = (long long)atomicAdd(counter, 1u);
    __syncthreads();
    const long long item = item_s;
    __syncthreads();
    if (item >= total) break;
    // instance: largest k with prefix[k] <= item (k < active)
    int lo = 0, hi = active - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (bd.prefix[mid] <= item) lo = mid;
      else hi = mid - 1;
    }
    const int i = bd.order[lo];
    if (done[i]) continue;
    const BatchInst<R> &I = inst[i];
    // plane geometry in the reference layout (lanes along the smaller stride)
    const int u = fr.ux, v = fr.uy;
    const bool u_first = I.stride[u] < I.stride[v];
    const int la = u_first ? u : v, lb = u_first ? v : u;
    const int na = I.n[la], nb = I.n[lb];
    const int nm = I.n[fr.m];
    const int p = fr.s > 0 ? 1 + q : nm - q;
    const long long base = (long long)p * I.stride[fr.m];
    const long long Sa = I.stride[la], Sb = I.stride[lb];
    const long long sm = fr.s * I.stride[fr.m];
    const int ua = fr.ux == la ? fr.sx : 0, ub = fr.ux == la ? 0 : fr.sx;
    const int va = fr.uy == la ? fr.sy : 0, vb = fr.uy == la ? 0 : fr.sy;
    const long long r = item - bd.prefix[lo];
    const int tiles_a = (na + 63) / 64;
    const int a0 = (int)(r % tiles_a) * 64, b = 1 + (int)(r / tiles_a);
    R *T = I.T;
    for (int idx = tx; idx < 3 * 66; idx += 64) {
      const int l_a = idx % 66, l_b = idx / 66;
      const int sa = a0 + l_a;
      if (sa <= na + 1) s[l_b][l_a] = T[base + (long long)sa * Sa + (long long)(b - 1 + l_b) * Sb];
    }
    __syncthreads();
    const int a = a0 + 1 + tx;
    R d = R(0);
    if (a <= na) {
      const long long t = base + (long long)a * Sa + (long long)b * Sb + sm;
      const int ca = tx + 1;
      R pv[9];
      pv[0] = s[1][ca];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ii = kRing[k][0], jj = kRing[k][1];
        pv[1 + k] = s[1 + ii * ub + jj * vb][ca + ii * ua + jj * va];
      }
      const R told = T[t];
      const R sl = I.V ? R(1) / I.V[t] : R(1) / I.speed;
      const int uu[3] = {fr.ux, fr.uy, fr.m};
      const R sg[3] = {R(fr.sx), R(fr.sy), R(fr.s)};
      const R hl[3] = {I.h[fr.ux], I.h[fr.uy], I.h[fr.m]};
      R N[6];
      if (I.th) {
        tilted_inverse_metric<R>(I.th[t], I.ph[t], I.ra[t], I.rb[t], N);
      } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) N[k] = I.Nc[k];
      }
      const LatticeN<R> L = lattice_of<R>(N, uu, sg, hl);
      const R tn = ellip_node_lattice<R>(told, pv, sl, L);
      if (tn != told) T[t] = tn;
      d = told - tn;
    }
    double vmax = (double)d;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
    if (tx == 0 && vmax > 0.0) atomicMax(&max_dt[i], (unsigned long long)__double_as_longlong(vmax));
    __syncthreads();  // s[][] is reused by the next item
  }
}

__global__ void batch_reset_kernel(unsigned int *counters, int nc, unsigned long long *max_dt,
                                   int ni) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc + ni; k += gridDim.x * blockDim.x) {
    if (k < nc) counters[k] = 0u;
    else max_dt[k - nc] = 0ull;
  }
}

template <typename R>
static BatchInst<R> make_inst(const fold3d_instance &x) {
  BatchInst<R> b;
  memset(&b, 0, sizeof b);
  b.T = (R *)x.T;
  b.V = (const R *)x.speed_field;
  b.th = (const R *)x.theta;
  b.ph = (const R *)x.phi;
  b.ra = (const R *)x.ra;
  b.rb = (const R *)x.rb;
  const Grid G(x.n[0], x.n[1], x.n[2]);
  for (int a = 0; a < 3; ++a) {
    b.stride[a] = G.stride[a];
    b.n[a] = x.n[a];
    b.h[a] = (R)x.h[a];
  }
  b.speed = (R)x.speed;
  double N[6];
  sym3_inverse(x.M, N);
  for (int k = 0; k < 6; ++k) b.Nc[k] = (R)N[k];
  return b;
}

// rows of 64 target nodes in one plane of direction d of instance x
static long long rows_of(const fold3d_instance &x, int d) {
  const Frame &fr = kFrames[d];
  const Grid G(x.n[0], x.n[1], x.n[2]);
  const bool u_first = G.stride[fr.ux] < G.stride[fr.uy];
  const int la = u_first ? fr.ux : fr.uy, lb = u_first ? fr.uy : fr.ux;
  return (long long)((x.n[la] + 63) / 64) * x.n[lb];
}

template <typename R>
int batch_solve(hipStream_t user_st, int count, const fold3d_instance *inst, int max_iters,
                double tol, int blocks, int use_graph, int *iters_out, double *max_dt_out,
                double *ms_out) {
  int dev = 0;
  F3D_HIP(hipGetDevice(&dev));
  // The batch runs on a stream of its own (the legacy default stream cannot be
  // captured), ordered after the caller's stream and before its later work.
  hipStream_t st = nullptr;
  hipEvent_t order_ev = nullptr;
  F3D_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  F3D_HIP(hipEventCreateWithFlags(&order_ev, hipEventDisableTiming));
  F3D_HIP(hipEventRecord(order_ev, user_st));
  F3D_HIP(hipStreamWaitEvent(st, order_ev, 0));
  bool capturing = false;
  // host tables
  std::vector<BatchInst<R>> hinst(count);
  for (int i = 0; i < count; ++i) hinst[i] = make_inst<R>(inst[i]);
  std::vector<int> order_all(6 * count);
  std::vector<long long> prefix_all(6 * (count + 1));
  std::vector<int> steps(6);
  for (int d = 0; d < 6; ++d) {
    const int m = kFrames[d].m;
    int *ord = &order_all[d * count];
    std::iota(ord, ord + count, 0);
    std::stable_sort(ord, ord + count, [&](int a, int b) { return inst[a].n[m] > inst[b].n[m]; });
    long long *pre = &prefix_all[d * (count + 1)];
    pre[0] = 0;
    for (int k = 0; k < count; ++k) pre[k + 1] = pre[k] + rows_of(inst[ord[k]], d);
    steps[d] = inst[ord[0]].n[m] - 1;
  }
  int launches = 0;
  for (int d = 0; d < 6; ++d) launches += steps[d];
  // device tables (freed at the end of the solve)
  BatchInst<R> *d_inst = nullptr;
  int *d_order = nullptr, *d_done = nullptr;
  long long *d_prefix = nullptr;
  unsigned int *d_counters = nullptr;
  unsigned long long *d_max = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FOLD3D_OK;
  auto fail = [&](hipError_t e, const char *what) {
    set_error("fold3d_solve_batched: %s failed: %s", what, hipGetErrorString(e));
    rc = e == hipErrorOutOfMemory ? FOLD3D_E_NOMEM : FOLD3D_E_HIP;
  };
#define BATCH_HIP(call)                         \
  do {                                          \
    hipError_t e_ = (call);                     \
    if (e_ != hipSuccess) {                     \
      fail(e_, #call);                          \
      goto cleanup;                             \
    }                                           \
  } while (0)
  {
    BATCH_HIP(hipMalloc(&d_inst, sizeof(BatchInst<R>) * count));
    BATCH_HIP(hipMalloc(&d_order, sizeof(int) * order_all.size()));
    BATCH_HIP(hipMalloc(&d_prefix, sizeof(long long) * prefix_all.size()));
    BATCH_HIP(hipMalloc(&d_done, sizeof(int) * count));
    BATCH_HIP(hipMalloc(&d_counters, sizeof(unsigned int) * launches));
    BATCH_HIP(hipMalloc(&d_max, sizeof(unsigned long long) * count));
    BATCH_HIP(hipMemcpyAsync(d_inst, hinst.data(), sizeof(BatchInst<R>) * count,
                             hipMemcpyHostToDevice, st));
    BATCH_HIP(hipMemcpyAsync(d_order, order_all.data(), sizeof(int) * order_all.size(),
                             hipMemcpyHostToDevice, st));
    BATCH_HIP(hipMemcpyAsync(d_prefix, prefix_all.data(), sizeof(long long) * prefix_all.size(),
                             hipMemcpyHostToDevice, st));
    std::vector<int> done(count, 0);
    BATCH_HIP(hipMemcpyAsync(d_done, done.data(), sizeof(int) * count, hipMemcpyHostToDevice, st));
    BATCH_HIP(hipStreamSynchronize(st));

    // one iteration as a launch sequence
    auto enqueue_iteration = [&](hipStream_t s) {
      batch_reset_kernel<<<(launches + count + 255) / 256, 256, 0, s>>>(d_counters, launches,
                                                                      d_max, count);
      int slot = 0;
      for (int d = 0; d < 6; ++d) {
        const int m = kFrames[d].m;
        const int *ord = &order_all[d * count];
        const long long *pre = &prefix_all[d * (count + 1)];
        BatchDir bd{d_order + d * count, d_prefix + d * (count + 1)};
        int active = count;
        for (int q = 0; q < steps[d]; ++q, ++slot) {
          while (active > 0 && inst[ord[active - 1]].n[m] - 1 <= q) --active;
          const long long total = pre[active];
          const int grid = (int)std::min<long long>(total, blocks);
          batch_plane_kernel<R><<<grid, 64, 0, s>>>(d_inst, bd, d, q, total, active,
                                                    d_counters + slot, d_done, d_max);
        }
      }
    };

    if (use_graph) {
      BATCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      capturing = true;
      enqueue_iteration(st);
      capturing = false;
      BATCH_HIP(hipStreamEndCapture(st, &graph));
      BATCH_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    BATCH_HIP(hipEventCreate(&e0));
    BATCH_HIP(hipEventCreate(&e1));
    BATCH_HIP(hipEventRecord(e0, st));
    std::vector<unsigned long long> hmax(count);
    std::vector<int> iters(count, 0);
    std::vector<double> last(count, -1.0);
    int remaining = count;
    for (int it = 0; it < max_iters && remaining > 0; ++it) {
      if (use_graph) BATCH_HIP(hipGraphLaunch(exec, st));
      else enqueue_iteration(st);
      BATCH_HIP(hipGetLastError());
      BATCH_HIP(hipMemcpyAsync(hmax.data(), d_max, sizeof(unsigned long long) * count,
                               hipMemcpyDeviceToHost, st));
      BATCH_HIP(hipStreamSynchronize(st));
      bool changed = false;
      for (int i = 0; i < count; ++i) {
        if (done[i]) continue;
        double v;
        memcpy(&v, &hmax[i], sizeof v);
        iters[i] = it + 1;
        last[i] = v;
        if (tol > 0 && v < tol) {
          done[i] = 1;
          --remaining;
          changed = true;
        }
      }
      if (changed)
        BATCH_HIP(hipMemcpyAsync(d_done, done.data(), sizeof(int) * count, hipMemcpyHostToDevice,
                                 st));
    }
    BATCH_HIP(hipEventRecord(e1, st));
    BATCH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    BATCH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_out) *ms_out = ms;
    for (int i = 0; i < count; ++i) {
      if (iters_out) iters_out[i] = iters[i];
      if (max_dt_out) max_dt_out[i] = last[i];
    }
    if (tol > 0 && remaining > 0) {
      set_error("fold3d_solve_batched: %d of %d instances not converged after %d iterations",
                remaining, count, max_iters);
      rc = FOLD3D_E_NOT_CONVERGED;
    }
  }
cleanup:
#undef BATCH_HIP
  if (capturing) {  // a failure inside the capture: end it before anything else
    hipGraph_t g2 = nullptr;
    (void)hipStreamEndCapture(st, &g2);
    if (g2) (void)hipGraphDestroy(g2);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  (void)hipFree(d_inst);
  (void)hipFree(d_order);
  (void)hipFree(d_prefix);
  (void)hipFree(d_done);
  (void)hipFree(d_counters);
  (void)hipFree(d_max);
  // later work on the caller's stream sees the solved fields
  (void)hipEventRecord(order_ev, st);
  (void)hipStreamWaitEvent(user_st, order_ev, 0);
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(order_ev);
  (void)hipStreamDestroy(st);
  return rc;
}

template int batch_solve<double>(hipStream_t, int, const fold3d_instance *, int, double, int, int,
                                 int *, double *, double *);
template int batch_solve<float>(hipStream_t, int, const fold3d_instance *, int, double, int, int,
                                int *, double *, double *);

}  // namespace fold3d
