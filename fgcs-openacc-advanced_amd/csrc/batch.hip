// batch.hip — many independent elliptic-model solves of uneven size in one
// stream of launches.
//
// Once per batch the slowness and the six lattice products of the inverse
// metric N = M^-1 of every instance are written to seven fields of its own in
// each sweep layout (tilt fields or a constant metric alike).  An iteration then runs, for all
// instances together, the single-instance metric engine's schedule of two
// moves: move T into the x-sweep layout (one launch for the whole batch), run
// the x-sweeps, move T into the y-sweep layout (y slowest), run the z- and the
// y-sweeps.  One launch per (sweep
// direction d, plane step q) serves every instance that still has a plane q in
// direction d; instances are sorted by their plane count along d, so those are
// a prefix of the sorted list.  A launch runs a fixed grid of persistent
// blocks (`groups` waves each, which share the candidate units of a row); a
// work item is one row of 64 target nodes of one instance, and a row table
// gives its instance and position in one load.  Blocks take items either from a
// per-launch device counter (`rows_per_pull` at a time) or by a static stride
// (block b takes items b, b + grid, ...); the batch autotuner picks.  The
// launches of one iteration (plus the counter and
// maximum resets) are captured once in a hipGraph and replayed; between
// replays the host reads each instance's max|dT| and retires converged
// instances through a device flag, so they cost only a flag test afterwards.
// Plane order and node update are those of the single-instance metric engine
// (same bits for per-node fields).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.h"
#include "ellip_local.h"
#include "metric_prep.h"
#include "solver.h"

namespace fold3d {

// Layouts of a batch: the reference layout Z, the x-sweep layout X (x
// slowest) and the layout Y (y slowest) in which the z- and y-sweeps run, as
// in the single-instance engines.
enum BatchLayout { kLayZ = 0, kLayX = 1, kLayY = 2 };

// Device view of one instance.  Scratch per instance: Y (T in layout X; T
// itself holds layout Y between iterations) and kMetricFields fields
// (the slowness and the lattice products of N = M^-1) in each sweep layout,
// MX and MY, made once per batch (metric_prep_kernel).
template <typename R>
struct BatchInst {
  R *T, *Y, *MX, *MY;
  long long count;
  long long stride[3][3];  // [layout][axis]
  int n[3];
};

// One row of 64 target nodes of a plane step: its instance, the row's first
// lane-axis index and its row-axis index.  Per direction the host lists the
// rows of one plane step, instance by instance in the sorted order, so that a
// block finds everything about its row in one load.
struct BatchRow {
  int inst, a0, b, pad;
};

// Lattice products of direction d from the metric fields, which hold the
// products P_q = (N_ab h_a) h_b (as lattice_map in the single-instance
// engine): L.X = sg * P_q.
struct LatticeIdx {
  int q[6];   // index of P_q (0..5) for A, B, C, E, Fx, Fy
  int sg[6];
};

// Rank of an axis in a batch layout's axis order (0: fastest).
constexpr int axis_rank(int lay, int axis) {
  const int *o = lay == kLayX ? kOrderX : lay == kLayY ? kOrderY : kOrderZ;
  return o[0] == axis ? 0 : o[1] == axis ? 1 : 2;
}

constexpr LatticeIdx lattice_idx(int d) {
  const Frame fr = kFrames[d];
  const int u[3] = {fr.ux, fr.uy, fr.m};
  const int sgn[3] = {fr.sx, fr.sy, fr.s};
  const int pairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  LatticeIdx L{};
  for (int e = 0; e < 6; ++e) {
    const int gi = u[pairs[e][0]], gj = u[pairs[e][1]];
    L.q[e] = gi == gj ? gi : (gi + gj == 1) ? 3 : (gi + gj == 2) ? 4 : 5;
    L.sg[e] = sgn[pairs[e][0]] * sgn[pairs[e][1]];
  }
  return L;
}

template <typename R, int G0, int kStride>
__device__ __forceinline__ void batch_units(R &best, const R p[9], R s2, const LatticeN<R> &L,
                                            R gap) {
  if constexpr (G0 < 8) {
    ellip_unit_lattice<R, G0, true>(best, p, s2, L, gap);
    batch_units<R, G0 + kStride, kStride>(best, p, s2, L, gap);
  }
}

template <typename R, int kGroups>
__device__ __forceinline__ void batch_dispatch_units(int grp, R &best, const R p[9], R s2,
                                                     const LatticeN<R> &L, R gap) {
  if constexpr (kGroups == 1) {
    batch_units<R, 0, 1>(best, p, s2, L, gap);
  } else if constexpr (kGroups == 2) {
    if (grp == 0) batch_units<R, 0, 2>(best, p, s2, L, gap);
    else batch_units<R, 1, 2>(best, p, s2, L, gap);
  } else {
    static_assert(kGroups == 4, "one, two or four waves per node row");
    switch (grp) {
      case 0: batch_units<R, 0, 4>(best, p, s2, L, gap); break;
      case 1: batch_units<R, 1, 4>(best, p, s2, L, gap); break;
      case 2: batch_units<R, 2, 4>(best, p, s2, L, gap); break;
      default: batch_units<R, 3, 4>(best, p, s2, L, gap); break;
    }
  }
}

// Largest |dT| of each instance in an iteration: kBatchMaxSlots words per
// instance, a block updating the one its index picks, so that the rows of a
// plane do not all meet on one word (the host takes the largest).
constexpr int kBatchMaxSlots = 32;

// One plane step q of direction kDir for every instance that has it:
// persistent blocks of kGroups waves pull rows of 64 target nodes from the
// launch's counter, `chunk` consecutive rows per pull.  The direction is a
// compile-time constant, so its frame and lattice indices take no registers.
template <typename R, int kGroups, int kDir>
__global__ __launch_bounds__(64 * kGroups) void batch_plane_kernel(
    const BatchInst<R> *__restrict__ inst, const BatchRow *__restrict__ rows, int q,
    long long total, int chunk,
    unsigned int *__restrict__ counter, const int *__restrict__ done,
    unsigned long long *__restrict__ max_dt) {
  __shared__ R s[3][64 + 2];
  __shared__ R part[kGroups][64];
  __shared__ long long item_s;
  constexpr Frame fr = kFrames[kDir];
  constexpr LatticeIdx LI = lattice_idx(kDir);
  // the x-sweeps run in layout X on the scratch copy Y, the others in layout Y on T
  constexpr int lay = fr.m == 0 ? kLayX : kLayY;
  constexpr bool on_T = fr.m != 0;
  // plane geometry: lanes along the in-plane axis of smaller stride, which the
  // layout's axis order fixes (strides grow strictly along it)
  constexpr bool kUFirst = axis_rank(lay, fr.ux) < axis_rank(lay, fr.uy);
  constexpr int la = kUFirst ? fr.ux : fr.uy, lb = kUFirst ? fr.uy : fr.ux;
  constexpr int ua = fr.ux == la ? fr.sx : 0, ub = fr.ux == la ? 0 : fr.sx;
  constexpr int va = fr.uy == la ? fr.sy : 0, vb = fr.uy == la ? 0 : fr.sy;
  const int tx = threadIdx.x;
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.y);
  // chunk > 0: pull `chunk` rows at a time from the launch's counter;
  // chunk == 0: static schedule, rows blockIdx.x, blockIdx.x + gridDim.x, ...
  // rows: the plane step's row table ([rows of one plane step, all
  // instances]), read with scalar loads; row indices are below 2^31.
  int item = chunk > 0 ? 0 : blockIdx.x, end = 0;
  for (;;) {
    if (chunk > 0) {
      if (item == end) {
        if (tx == 0 && grp == 0) item_s = (long long)atomicAdd(counter, (unsigned)chunk);
        __syncthreads();
        item = __builtin_amdgcn_readfirstlane((int)item_s);
        __syncthreads();
        if (item >= total) break;
        end = item + chunk < total ? item + chunk : (int)total;
      }
    } else if (item >= total) {
      break;
    }
    const BatchRow rw = rows[item];
    item += chunk > 0 ? 1 : gridDim.x;
    const int i = __builtin_amdgcn_readfirstlane(rw.inst);
    const BatchInst<R> &I = inst[i];
    const long long *st = I.stride[lay];
    {  // the row's instance data as one group of scalar loads (one wait, not
       // one per use)
      const int dn = done[i];
      asm volatile("" ::"s"(dn), "s"(st[la]), "s"(st[lb]), "s"(st[fr.m]), "s"(I.n[la]),
                   "s"(I.n[fr.m]), "s"(I.count), "s"(on_T ? I.T : I.Y),
                   "s"(lay == kLayX ? I.MX : I.MY));
      if (dn) continue;  // uniform over the block
    }
    const int na = I.n[la];
    const int nm = I.n[fr.m];
    const int p = fr.s > 0 ? 1 + q : nm - q;
    const long long base = (long long)p * st[fr.m];
    const long long Sa = st[la], Sb = st[lb];
    const long long sm = fr.s * st[fr.m];
    const int a0 = __builtin_amdgcn_readfirstlane(rw.a0), b = __builtin_amdgcn_readfirstlane(rw.b);
    R *T = on_T ? I.T : I.Y;
    const R *M = lay == kLayX ? I.MX : I.MY;
    // all loads of the row in flight together (RowStage); a lane outside the
    // plane reads the row's first target node instead, and drops it
    const int a = a0 + 1 + tx;
    const bool act = a <= na;
    const long long t = base + (long long)a * Sa + (long long)b * Sb + sm;
    const long long tl = act ? t : t + (long long)(a0 + 1 - a) * Sa;
    const R tload = T[tl], sload = M[tl];
    R Mv[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) Mv[e] = M[(1 + LI.q[e]) * I.count + tl];
    RowStage<R, 64 * kGroups> rs;
    rs.load(T, base, Sa, Sb, a0, b, na, grp * 64 + tx);
    rs.store(s, grp * 64 + tx);
    const R told = act ? tload : R(0), sl = act ? sload : R(1);
    R Lv[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) Lv[e] = LI.sg[e] > 0 ? Mv[e] : -Mv[e];  // sg * P_q
    LatticeN<R> L;
    L.A = Lv[0];
    L.B = Lv[1];
    L.C = Lv[2];
    L.E = Lv[3];
    L.Fx = Lv[4];
    L.Fy = Lv[5];
    __syncthreads();
    R best = told;
    if (act) {
      // the patch (uniform offsets of +x^ and +y^ within the staged rows)
      const R *c0 = &s[1][tx + 1];
      const int ox = ub * 66 + ua, oy = vb * 66 + va;
      R pv[9];
      pv[0] = c0[0];
#pragma unroll
      for (int k = 0; k < 8; ++k) pv[1 + k] = c0[kRing[k][0] * ox + kRing[k][1] * oy];
      ellip_vertices_lattice<R>(best, pv, sl, L);
      batch_dispatch_units<R, kGroups>(grp, best, pv, sl * sl, L, plane_gap(L, sl));
    }
    part[grp][tx] = best;
    __syncthreads();
    if (grp == 0) {
      R d = R(0);
      if (act) {
        R mn = part[0][tx];
#pragma unroll
        for (int k = 1; k < kGroups; ++k) mn = rmin(part[k][tx], mn);
        if (mn != told) T[t] = mn;
        d = told - mn;
      }
      const double vmax = wave_max_dt(d);
      if (tx == 0 && vmax > 0.0)  // the instance's word for this block (see kBatchMaxSlots)
        atomicMax(&max_dt[(long long)i * kBatchMaxSlots + blockIdx.x % kBatchMaxSlots],
                  (unsigned long long)__double_as_longlong(vmax));
    }
    __syncthreads();  // s[][] and part[][] are reused by the next item
  }
}

// Arguments of one plane-step launch, and the launch by direction.
template <typename R>
struct BatchLaunch {
  const BatchInst<R> *inst;
  const BatchRow *rows;
  int q;
  long long total;
  int chunk;
  unsigned int *counter;
  const int *done;
  unsigned long long *max_dt;
};
template <typename R, int kGroups, int kDir>
static void launch_batch_plane_dir(int grid, hipStream_t s, const BatchLaunch<R> &a) {
  batch_plane_kernel<R, kGroups, kDir><<<grid, dim3(64, kGroups), 0, s>>>(
      a.inst, a.rows, a.q, a.total, a.chunk, a.counter, a.done, a.max_dt);
}
template <typename R, int kGroups>
static void launch_batch_plane(int d, int grid, hipStream_t s, const BatchLaunch<R> &a) {
  switch (d) {
    case 0: launch_batch_plane_dir<R, kGroups, 0>(grid, s, a); break;
    case 1: launch_batch_plane_dir<R, kGroups, 1>(grid, s, a); break;
    case 2: launch_batch_plane_dir<R, kGroups, 2>(grid, s, a); break;
    case 3: launch_batch_plane_dir<R, kGroups, 3>(grid, s, a); break;
    case 4: launch_batch_plane_dir<R, kGroups, 4>(grid, s, a); break;
    default: launch_batch_plane_dir<R, kGroups, 5>(grid, s, a); break;
  }
}

// Moves `fields` consecutive fields of every active instance between two
// layouts: block = one 32x32 tile of one (x, y) plane of one instance.
constexpr int kBTile = 32, kBTileRows = 8;

template <typename R>
struct PermuteJob {
  int which;          // 0: T -> Y, 1: Y -> T
  int lin, lout;      // layouts
  int fields;
  int all;            // also the instances that are done
};

// One tile of a layout move: its instance, its 32x32 column (bx, by) and its
// plane z, listed by the host so that a block finds them in one load.
struct BatchTile {
  int inst, bx, by, z;
};

template <typename R>
__global__ __launch_bounds__(kBTile *kBTileRows) void batch_permute_kernel(
    const BatchInst<R> *__restrict__ inst, const BatchTile *__restrict__ tiles,
    PermuteJob<R> job, const int *__restrict__ done) {
  __shared__ R tile[kBTile][kBTile + 1];  // [y - y0][x - x0]
  const BatchTile tl = tiles[blockIdx.x];
  const int lo = tl.inst;
  if (!job.all && done[lo]) return;
  const BatchInst<R> &I = inst[lo];
  const int Hx = I.n[0] + 2, Hy = I.n[1] + 2;
  const int bx = tl.bx, by = tl.by;
  const long long z = tl.z;
  const long long *is = I.stride[job.lin], *os = I.stride[job.lout];
  const bool in_x = is[0] == 1, out_x = os[0] == 1;
  const R *src = job.which == 0 ? I.T : I.Y;
  R *dst = job.which == 0 ? I.Y : I.T;
  const int x0 = bx * kBTile, y0 = by * kBTile;
  const int tx = threadIdx.x;
  for (int f = 0; f < job.fields; ++f) {
    const R *in = src + f * I.count + z * is[2];
    R *out = dst + f * I.count + z * os[2];
#pragma unroll
    for (int rr = threadIdx.y; rr < kBTile; rr += kBTileRows) {
      const int x = x0 + (in_x ? tx : rr), y = y0 + (in_x ? rr : tx);
      if (x < Hx && y < Hy) tile[y - y0][x - x0] = in[x * is[0] + y * is[1]];
    }
    __syncthreads();
#pragma unroll
    for (int rr = threadIdx.y; rr < kBTile; rr += kBTileRows) {
      const int x = x0 + (out_x ? tx : rr), y = y0 + (out_x ? rr : tx);
      if (x < Hx && y < Hy) out[x * os[0] + y * os[1]] = tile[y - y0][x - x0];
    }
    __syncthreads();
  }
}

__global__ void batch_reset_kernel(unsigned int *counters, int nc, unsigned long long *max_dt,
                                   int ni) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc + ni; k += gridDim.x * blockDim.x) {
    if (k < nc) counters[k] = 0u;
    else max_dt[k - nc] = 0ull;
  }
}

// Rows of 64 target nodes in one plane of direction d of instance x, in the
// layout that direction runs in (lanes along the in-plane axis of smaller
// stride, as the kernel picks them): `tiles` per line, `lines` lines.
static void row_shape(const fold3d_instance &x, int d, int &tiles, int &lines) {
  const Frame &fr = kFrames[d];
  const int *order = fr.m == 0 ? kOrderX : fr.m == 2 ? kOrderZ : kOrderY;
  const Grid G(x.n[0], x.n[1], x.n[2], order);
  const bool u_first = G.stride[fr.ux] < G.stride[fr.uy];
  const int la = u_first ? fr.ux : fr.uy, lb = u_first ? fr.uy : fr.ux;
  tiles = (x.n[la] + 63) / 64;
  lines = x.n[lb];
}
static long long rows_of(const fold3d_instance &x, int d) {
  int tiles, lines;
  row_shape(x, d, tiles, lines);
  return (long long)tiles * lines;
}

size_t batch_scratch_bytes(int count, const fold3d_instance *inst, int precision) {
  const size_t r = precision == FOLD3D_F64 ? 8 : 4;
  const int fields = 1 + 2 * kMetricFields;
  size_t total = 0;
  for (int i = 0; i < count; ++i) {
    const Grid G(inst[i].n[0], inst[i].n[1], inst[i].n[2]);
    total += (size_t)G.count * r * fields;
    total = (total + 255) & ~(size_t)255;
  }
  return total;
}

template <typename R>
int batch_solve(hipStream_t user_st, int count, const fold3d_instance *inst, int max_iters,
                double tol, int blocks, int groups, int rows_per_pull, int use_graph,
                void *scratch, int *iters_out, double *max_dt_out, double *ms_out) {
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
  std::vector<long long> tiles_prefix(count + 1, 0);
  std::vector<int> tiles_nx(count), tiles_ny(count);
  {
    char *p = (char *)scratch;
    for (int i = 0; i < count; ++i) {
      const fold3d_instance &x = inst[i];
      BatchInst<R> &b = hinst[i];
      memset(&b, 0, sizeof b);
      const Grid GZ(x.n[0], x.n[1], x.n[2], kOrderZ), GX = GZ.with_order(kOrderX),
          GY = GZ.with_order(kOrderY);
      b.T = (R *)x.T;
      b.count = GZ.count;
      b.Y = (R *)p;
      b.MX = b.Y + GZ.count;
      b.MY = b.MX + kMetricFields * GZ.count;
      p += (size_t)GZ.count * sizeof(R) * (1 + 2 * kMetricFields);
      p = (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255);
      for (int a = 0; a < 3; ++a) {
        b.stride[kLayZ][a] = GZ.stride[a];
        b.stride[kLayX][a] = GX.stride[a];
        b.stride[kLayY][a] = GY.stride[a];
        b.n[a] = x.n[a];
      }
      const long long tiles = (long long)((x.n[0] + 2 + kBTile - 1) / kBTile) *
                              ((x.n[1] + 2 + kBTile - 1) / kBTile) * (x.n[2] + 2);
      tiles_prefix[i + 1] = tiles_prefix[i] + tiles;
      tiles_nx[i] = (x.n[0] + 2 + kBTile - 1) / kBTile;
      tiles_ny[i] = (x.n[1] + 2 + kBTile - 1) / kBTile;
    }
  }
  const long long total_tiles = tiles_prefix[count];
  std::vector<BatchTile> tile_all(total_tiles);
  for (int i = 0; i < count; ++i) {
    BatchTile *t = &tile_all[tiles_prefix[i]];
    const int planes = inst[i].n[2] + 2;
    for (int z = 0; z < planes; ++z)
      for (int by = 0; by < tiles_ny[i]; ++by)
        for (int bx = 0; bx < tiles_nx[i]; ++bx) *t++ = BatchTile{i, bx, by, z};
  }
  std::vector<int> order_all(6 * count);
  std::vector<long long> prefix_all(6 * (count + 1));
  std::vector<int> steps(6);
  std::vector<long long> row_off(7, 0);
  for (int d = 0; d < 6; ++d) {
    const int m = kFrames[d].m;
    int *ord = &order_all[d * count];
    std::iota(ord, ord + count, 0);
    std::stable_sort(ord, ord + count, [&](int a, int b) { return inst[a].n[m] > inst[b].n[m]; });
    long long *pre = &prefix_all[d * (count + 1)];
    pre[0] = 0;
    for (int k = 0; k < count; ++k) pre[k + 1] = pre[k] + rows_of(inst[ord[k]], d);
    steps[d] = inst[ord[0]].n[m] - 1;
    row_off[d + 1] = row_off[d] + pre[count];
  }
  std::vector<BatchRow> row_all(row_off[6]);
  for (int d = 0; d < 6; ++d) {
    const int *ord = &order_all[d * count];
    const long long *pre = &prefix_all[d * (count + 1)];
    BatchRow *row = &row_all[row_off[d]];
    for (int k = 0; k < count; ++k) {
      int tiles, lines;
      row_shape(inst[ord[k]], d, tiles, lines);
      for (long long r = pre[k]; r < pre[k + 1]; ++r) {
        const long long q = r - pre[k];
        row[r] = BatchRow{ord[k], (int)(q % tiles) * 64, 1 + (int)(q / tiles), 0};
      }
    }
  }
  int launches = 0;
  for (int d = 0; d < 6; ++d) launches += steps[d];
  // device tables (freed at the end of the solve)
  BatchInst<R> *d_inst = nullptr;
  int *d_done = nullptr;
  BatchRow *d_rows = nullptr;
  BatchTile *d_tiles = nullptr;
  unsigned int *d_counters = nullptr;
  unsigned long long *d_max = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
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
    BATCH_HIP(hipMalloc(&d_done, sizeof(int) * count));
    BATCH_HIP(hipMalloc(&d_counters, sizeof(unsigned int) * launches));
    BATCH_HIP(hipMalloc(&d_max, sizeof(unsigned long long) * count * kBatchMaxSlots));
    BATCH_HIP(hipMalloc(&d_tiles, sizeof(BatchTile) * tile_all.size()));
    BATCH_HIP(hipMalloc(&d_rows, sizeof(BatchRow) * row_all.size()));
    BATCH_HIP(hipMemcpyAsync(d_rows, row_all.data(), sizeof(BatchRow) * row_all.size(),
                             hipMemcpyHostToDevice, st));
    BATCH_HIP(hipMemcpyAsync(d_tiles, tile_all.data(), sizeof(BatchTile) * tile_all.size(),
                             hipMemcpyHostToDevice, st));
    BATCH_HIP(hipMemcpyAsync(d_inst, hinst.data(), sizeof(BatchInst<R>) * count,
                             hipMemcpyHostToDevice, st));
    std::vector<int> done(count, 0);
    BATCH_HIP(hipMemcpyAsync(d_done, done.data(), sizeof(int) * count, hipMemcpyHostToDevice, st));
    // slowness and lattice products of every instance in both sweep
    // layouts, once per batch
    for (int i = 0; i < count; ++i) {
      const fold3d_instance &x = inst[i];
      double N[6];
      sym3_inverse(x.M, N);
      launch_metric_prep<R>(st, hinst[i].count, x.n, hinst[i].stride[kLayZ],
                            hinst[i].stride[kLayX], hinst[i].stride[kLayY],
                            (const R *)x.speed_field, (R)x.speed, (const R *)x.theta,
                            (const R *)x.phi, (const R *)x.ra, (const R *)x.rb, N, x.h,
                            hinst[i].MX, hinst[i].MY);
    }
    BATCH_HIP(hipGetLastError());
    BATCH_HIP(hipStreamSynchronize(st));

    // One iteration as a launch sequence, two moves of T: into the x-sweep
    // layout in the scratch field Y, the x-sweeps there, back into the
    // caller's T in the y-sweep layout (y slowest), the z- and y-sweeps there.
    // Between iterations T stays in the y-sweep layout (the first iteration
    // starts from the reference layout); after the last, every instance's T
    // goes back to the reference layout.
    auto enqueue_iteration = [&](hipStream_t s, bool first) {
      batch_reset_kernel<<<(launches + count * kBatchMaxSlots + 255) / 256, 256, 0, s>>>(
          d_counters, launches, d_max, count * kBatchMaxSlots);
      int slot = 0;
      auto permute = [&](int which, int lin, int lout, int fields) {
        PermuteJob<R> job{which, lin, lout, fields, 0};
        batch_permute_kernel<R><<<(unsigned)total_tiles, dim3(kBTile, kBTileRows), 0, s>>>(
            d_inst, d_tiles, job, d_done);
      };
      // (the x-sweeps run in layout X on the scratch copy, the others in
      // layout Y on the caller's T: batch_plane_kernel derives both from d)
      auto direction = [&](int d) {
        const int m = kFrames[d].m;
        const int *ord = &order_all[d * count];
        const long long *pre = &prefix_all[d * (count + 1)];
        const BatchRow *rows_d = d_rows + row_off[d];
        int active = count;
        for (int q = 0; q < steps[d]; ++q, ++slot) {
          while (active > 0 && inst[ord[active - 1]].n[m] - 1 <= q) --active;
          const long long total = pre[active];
          // rows per pull: enough pulls per block to balance, few enough atomics
          // rows_per_pull < 0: static schedule; 0: automatic pull size
          const int chunk =
              rows_per_pull < 0   ? 0
              : rows_per_pull > 0 ? rows_per_pull
                                  : (int)std::max<long long>(
                                        1, std::min<long long>(8, total / (4 * blocks)));
          const int grid =
              (int)std::min<long long>(chunk > 0 ? (total + chunk - 1) / chunk : total, blocks);
          const BatchLaunch<R> bl{d_inst, rows_d, q, total, chunk, d_counters + slot, d_done, d_max};
          if (groups == 1) launch_batch_plane<R, 1>(d, grid, s, bl);
          else if (groups == 2) launch_batch_plane<R, 2>(d, grid, s, bl);
          else launch_batch_plane<R, 4>(d, grid, s, bl);
        }
      };
      permute(0, first ? kLayZ : kLayY, kLayX, 1);
      direction(FOLD3D_X_DOWN);
      direction(FOLD3D_X_UP);
      permute(1, kLayX, kLayY, 1);
      direction(FOLD3D_K_UP);
      direction(FOLD3D_K_DOWN);
      direction(FOLD3D_Y_UP);
      direction(FOLD3D_Y_DOWN);
    };

    if (use_graph) {  // the first iteration and the others
      for (int g = 0; g < 2; ++g) {
        BATCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        capturing = true;
        enqueue_iteration(st, g == 0);
        capturing = false;
        BATCH_HIP(hipStreamEndCapture(st, &graph[g]));
        BATCH_HIP(hipGraphInstantiate(&exec[g], graph[g], nullptr, nullptr, 0));
      }
    }
    BATCH_HIP(hipEventCreate(&e0));
    BATCH_HIP(hipEventCreate(&e1));
    BATCH_HIP(hipEventRecord(e0, st));
    std::vector<unsigned long long> hmax((size_t)count * kBatchMaxSlots);
    std::vector<int> iters(count, 0);
    std::vector<double> last(count, -1.0);
    int remaining = count;
    int ran = 0;  // iterations launched
    for (int it = 0; it < max_iters && remaining > 0; ++it, ++ran) {
      if (use_graph) BATCH_HIP(hipGraphLaunch(exec[it == 0 ? 0 : 1], st));
      else enqueue_iteration(st, it == 0);
      BATCH_HIP(hipGetLastError());
      BATCH_HIP(hipMemcpyAsync(hmax.data(), d_max, sizeof(unsigned long long) * hmax.size(),
                               hipMemcpyDeviceToHost, st));
      BATCH_HIP(hipStreamSynchronize(st));
      bool changed = false;
      for (int i = 0; i < count; ++i) {
        if (done[i]) continue;
        unsigned long long m = 0;  // largest of the instance's words
        for (int k = 0; k < kBatchMaxSlots; ++k)
          m = std::max(m, hmax[(size_t)i * kBatchMaxSlots + k]);
        double v;
        memcpy(&v, &m, sizeof v);
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
    if (ran > 0) {  // every instance's T back into the reference layout, through Y
      const PermuteJob<R> back{0, kLayY, kLayZ, 1, 1}, home{1, kLayZ, kLayZ, 1, 1};
      batch_permute_kernel<R><<<(unsigned)total_tiles, dim3(kBTile, kBTileRows), 0, st>>>(
          d_inst, d_tiles, back, d_done);
      batch_permute_kernel<R><<<(unsigned)total_tiles, dim3(kBTile, kBTileRows), 0, st>>>(
          d_inst, d_tiles, home, d_done);
      BATCH_HIP(hipGetLastError());
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
  for (int g = 0; g < 2; ++g) {
    if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
    if (graph[g]) (void)hipGraphDestroy(graph[g]);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  (void)hipFree(d_inst);
  (void)hipFree(d_done);
  (void)hipFree(d_counters);
  (void)hipFree(d_max);
  (void)hipFree(d_tiles);
  (void)hipFree(d_rows);
  // later work on the caller's stream sees the solved fields
  (void)hipEventRecord(order_ev, st);
  (void)hipStreamWaitEvent(user_st, order_ev, 0);
  (void)hipStreamSynchronize(st);
  (void)hipEventDestroy(order_ev);
  (void)hipStreamDestroy(st);
  return rc;
}

template int batch_solve<double>(hipStream_t, int, const fold3d_instance *, int, double, int, int,
                                 int, int, void *, int *, double *, double *);
template int batch_solve<float>(hipStream_t, int, const fold3d_instance *, int, double, int, int,
                                int, int, void *, int *, double *, double *);

}  // namespace fold3d
