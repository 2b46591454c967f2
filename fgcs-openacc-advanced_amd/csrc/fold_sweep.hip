// fold_sweep.hip — plane-step kernels and the sweep driver of the fold-model
// PMM solve on gfx950, behind the C-ABI of include/fold3d.h.
//
// Reference: Stencil-OpenACC/AN_ISO_1_OpenACC.c — plane kernels :180-1092,
// driver :1150-1241 (6 directions x (n-1) planes per iteration, 8 iterations).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "common.h"
#include "fold_fast.h"
#include "fold_local.h"
#include "solver.h"

namespace fold3d {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

}  // namespace fold3d

extern "C" const char *fold3d_last_error(void) { return fold3d::g_err; }

// for the host C parts of the library (io.c)
extern "C" void fold3d_set_error_text(const char *msg) { fold3d::set_error("%s", msg); }

namespace fold3d {

// ---------------------------------------------------------------------------
// Plane-step kernel: one thread per node of the target plane.  Lanes run along
// the in-plane axis with the smallest stride (x for y- and z-sweeps, which is
// unit stride and coalesced; y for x-sweeps).  Blocks are 64 x 4: one
// wavefront per row.
// ---------------------------------------------------------------------------
constexpr int kLaneX = 64;
constexpr int kRowsY = 4;

template <typename R, bool kReduce>
__global__ __launch_bounds__(kLaneX *kRowsY) void fold_plane_kernel(
    R *__restrict__ T, PlaneGeom g, FoldFrame<R> f,
    unsigned long long *__restrict__ max_dt) {
  const int a = 1 + blockIdx.x * kLaneX + threadIdx.x;
  const int b = 1 + blockIdx.y * kRowsY + threadIdx.y;
  R d = R(0);
  if (a <= g.na && b <= g.nb) {
    const long long c = g.base + a * g.Sa + b * g.Sb;
    Patch<R> p;
    p.c = fabs(T[c]);
    p.px = fabs(T[c + g.su]);
    p.py = fabs(T[c + g.sv]);
    p.mx = fabs(T[c - g.su]);
    p.my = fabs(T[c - g.sv]);
    p.pxpy = fabs(T[c + g.su + g.sv]);
    p.pxmy = fabs(T[c + g.su - g.sv]);
    p.mxpy = fabs(T[c - g.su + g.sv]);
    p.mxmy = fabs(T[c - g.su - g.sv]);
    const R told = T[c + g.sm];
    const R tn = fold_node_update(told, p, f);
    T[c + g.sm] = tn;
    if (kReduce) d = fabs(tn - told);
  }
  if (kReduce) {
    // wavefront max, then one atomic per wave (|dT| >= 0, so the IEEE bit
    // pattern orders like an unsigned integer)
    const double v = wave_max_dt(d);
    if ((threadIdx.x & 63) == 0 && v > 0.0)
      max_dt_update(max_dt, v);
  }
}

PlaneGeom plane_geom(const Grid &G, int dir, int plane) {
  const Frame &fr = kFrames[dir];
  // lane axis: the in-plane axis with the smaller stride
  const int u = fr.ux, v = fr.uy;
  const bool u_first = G.stride[u] < G.stride[v];
  const int la = u_first ? u : v, lb = u_first ? v : u;
  PlaneGeom g;
  g.base = (long long)plane * G.stride[fr.m];
  g.Sa = G.stride[la];
  g.Sb = G.stride[lb];
  g.sm = fr.s * G.stride[fr.m];
  g.su = fr.sx * G.stride[fr.ux];
  g.sv = fr.sy * G.stride[fr.uy];
  g.na = G.n[la];
  g.nb = G.n[lb];
  g.ua = fr.ux == la ? fr.sx : 0;
  g.ub = fr.ux == la ? 0 : fr.sx;
  g.va = fr.uy == la ? fr.sy : 0;
  g.vb = fr.uy == la ? 0 : fr.sy;
  return g;
}

template <typename R>
static void launch_plane(hipStream_t st, R *T, const PlaneGeom &g,
                         const FoldFrame<R> &f, unsigned long long *max_dt) {
  dim3 block(kLaneX, kRowsY);
  dim3 grid((g.na + kLaneX - 1) / kLaneX, (g.nb + kRowsY - 1) / kRowsY);
  if (max_dt)
    fold_plane_kernel<R, true><<<grid, block, 0, st>>>(T, g, f, max_dt);
  else
    fold_plane_kernel<R, false><<<grid, block, 0, st>>>(T, g, f, nullptr);
}

// One direction: the reference's plane order (forward 1..n-1, backward n..2).
template <typename R>
static void fold_direction(hipStream_t st, R *T, const Grid &G, int dir,
                           const FoldFrame<R> &f, unsigned long long *max_dt) {
  const Frame &fr = kFrames[dir];
  const int n = G.n[fr.m];
  if (fr.s > 0)
    for (int p = 1; p < n; ++p) launch_plane(st, T, plane_geom(G, dir, p), f, max_dt);
  else
    for (int p = n; p > 1; --p) launch_plane(st, T, plane_geom(G, dir, p), f, max_dt);
}

template <typename R>
int fold_sweeps_plane_engine(hipStream_t st, R *T, const Grid &G,
                             const fold3d_model &m, const double h[3],
                             int max_iters, double tol,
                             unsigned long long *d_max, int *iters_out,
                             double *max_dt_out) {
  FoldFrame<R> frames[6];
  for (int d = 0; d < 6; ++d) frames[d] = fold_frame<R>(d, m.F, m.a, h);
  int it = 0;
  double last = -1.0;
  for (; it < max_iters;) {
    if (tol > 0) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    for (int d = 0; d < 6; ++d)
      fold_direction<R>(st, T, G, d, frames[d], tol > 0 ? d_max : nullptr);
    F3D_HIP(hipGetLastError());
    ++it;
    if (tol > 0) {
      F3D_HIP(max_dt_read(st, d_max, &last));
      if (last < tol) break;
    }
  }
  *iters_out = it;
  *max_dt_out = last;
  return FOLD3D_OK;
}

template int fold_sweeps_plane_engine<double>(hipStream_t, double *, const Grid &,
                                              const fold3d_model &, const double *,
                                              int, double, unsigned long long *,
                                              int *, double *);
template int fold_sweeps_plane_engine<float>(hipStream_t, float *, const Grid &,
                                             const fold3d_model &, const double *,
                                             int, double, unsigned long long *,
                                             int *, double *);

int fold_plane_step_f64(hipStream_t st, double *T, const Grid &G, int dir,
                        int plane, const FoldFrame<double> &f) {
  launch_plane(st, T, plane_geom(G, dir, plane), f, nullptr);
  F3D_HIP(hipGetLastError());
  return FOLD3D_OK;
}

// ---------------------------------------------------------------------------
// Hoisted-constant engine: same plane loop, node update from fold_fast.h.
// C points at this direction's FoldFast in device memory; the two sign
// variants sit in separate branches so a wave of one sign runs one of them
// with its constants in scalar registers.
// ---------------------------------------------------------------------------
template <typename R, bool kReduce>
__global__ __launch_bounds__(kLaneX *kRowsY) void fold_plane_fast_kernel(
    R *__restrict__ T, PlaneGeom g, const FoldFast<R> *__restrict__ C,
    unsigned long long *__restrict__ max_dt) {
  const int a = 1 + blockIdx.x * kLaneX + threadIdx.x;
  const int b = 1 + blockIdx.y * kRowsY + threadIdx.y;
  R d = R(0);
  if (a <= g.na && b <= g.nb) {
    const long long c = g.base + a * g.Sa + b * g.Sb;
    Patch<R> p;
    p.c = fabs(T[c]);
    p.px = fabs(T[c + g.su]);
    p.py = fabs(T[c + g.sv]);
    p.mx = fabs(T[c - g.su]);
    p.my = fabs(T[c - g.sv]);
    p.pxpy = fabs(T[c + g.su + g.sv]);
    p.pxmy = fabs(T[c + g.su - g.sv]);
    p.mxpy = fabs(T[c - g.su + g.sv]);
    p.mxmy = fabs(T[c - g.su - g.sv]);
    const R told = T[c + g.sm];
    R tn;
    if (told < R(0))
      tn = fold_node_fast(told, p, C->sgn[1]);
    else
      tn = fold_node_fast(told, p, C->sgn[0]);
    if (tn != told) T[c + g.sm] = tn;
    if (kReduce) d = fabs(tn - told);
  }
  if (kReduce) {
    const double v = wave_max_dt(d);
    if ((threadIdx.x & 63) == 0 && v > 0.0)
      max_dt_update(max_dt, v);
  }
}

template <typename R>
void build_fold_fast(int dir, const fold3d_model &m, const double h[3], FoldFast<R> &out) {
  // the frame in R, as the literal kernel receives it, then everything in R
  fold_fast_consts<R>(dir, m.F, m.a, h, out);
}
template void build_fold_fast<double>(int, const fold3d_model &, const double *, FoldFast<double> &);
template void build_fold_fast<float>(int, const fold3d_model &, const double *, FoldFast<float> &);

template <typename R>
int fold_sweeps_fast_engine(hipStream_t st, R *T, const Grid &G, const FoldFast<R> *d_consts,
                            int max_iters, double tol, unsigned long long *d_max,
                            int *iters_out, double *max_dt_out) {
  int it = 0;
  double last = -1.0;
  const dim3 block(kLaneX, kRowsY);
  for (; it < max_iters;) {
    if (tol > 0) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    for (int d = 0; d < 6; ++d) {
      const Frame &fr = kFrames[d];
      const int n = G.n[fr.m];
      for (int q = 0; q < n - 1; ++q) {
        const int p = fr.s > 0 ? 1 + q : n - q;
        const PlaneGeom g = plane_geom(G, d, p);
        const dim3 grid((g.na + kLaneX - 1) / kLaneX, (g.nb + kRowsY - 1) / kRowsY);
        if (tol > 0)
          fold_plane_fast_kernel<R, true><<<grid, block, 0, st>>>(T, g, d_consts + d, d_max);
        else
          fold_plane_fast_kernel<R, false><<<grid, block, 0, st>>>(T, g, d_consts + d, nullptr);
      }
    }
    F3D_HIP(hipGetLastError());
    ++it;
    if (tol > 0) {
      F3D_HIP(max_dt_read(st, d_max, &last));
      if (last < tol) break;
    }
  }
  *iters_out = it;
  *max_dt_out = last;
  return FOLD3D_OK;
}
template int fold_sweeps_fast_engine<double>(hipStream_t, double *, const Grid &,
                                             const FoldFast<double> *, int, double,
                                             unsigned long long *, int *, double *);
template int fold_sweeps_fast_engine<float>(hipStream_t, float *, const Grid &,
                                            const FoldFast<float> *, int, double,
                                            unsigned long long *, int *, double *);

// ---------------------------------------------------------------------------
// Tiled engine.  A block owns a kTileA x kTileB tile of the target plane; it
// stages the source-plane tile plus a one-node halo, as magnitudes, in LDS
// (each source node is read from HBM/L2 once per block instead of nine times
// per node) and reads the ten-value patch of every node from there.  Rows of
// the tile run along the plane's unit-stride axis in every layout.
// ---------------------------------------------------------------------------
constexpr int kTileA = 64;
constexpr int kTileB = 4;

template <typename R, bool kReduce>
__global__ __launch_bounds__(kTileA *kTileB) void fold_tile_kernel(
    R *__restrict__ T, PlaneGeom g, const FoldFast<R> *__restrict__ C,
    unsigned long long *__restrict__ max_dt) {
  __shared__ R s[kTileB + 2][kTileA + 2];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int a0 = blockIdx.x * kTileA, b0 = blockIdx.y * kTileB;
  const int a = a0 + 1 + tx, b = b0 + 1 + ty;
  const bool active = a <= g.na && b <= g.nb;
  const long long c = g.base + (long long)a * g.Sa + (long long)b * g.Sb;
  const R told = active ? T[c + g.sm] : R(0);
  for (int idx = ty * kTileA + tx; idx < (kTileA + 2) * (kTileB + 2); idx += kTileA * kTileB) {
    const int la = idx % (kTileA + 2), lb = idx / (kTileA + 2);
    const int sa = a0 + la, sb = b0 + lb;
    if (sa <= g.na + 1 && sb <= g.nb + 1)
      s[lb][la] = fabs(T[g.base + (long long)sa * g.Sa + (long long)sb * g.Sb]);
  }
  __syncthreads();
  R d = R(0);
  if (active) {
    const int ca = tx + 1, cb = ty + 1;
    const int ua = g.ua, ub = g.ub, va = g.va, vb = g.vb;
    Patch<R> p;
    p.c = s[cb][ca];
    p.px = s[cb + ub][ca + ua];
    p.py = s[cb + vb][ca + va];
    p.mx = s[cb - ub][ca - ua];
    p.my = s[cb - vb][ca - va];
    p.pxpy = s[cb + ub + vb][ca + ua + va];
    p.pxmy = s[cb + ub - vb][ca + ua - va];
    p.mxpy = s[cb - ub + vb][ca - ua + va];
    p.mxmy = s[cb - ub - vb][ca - ua - va];
    R tn;
    if (told < R(0))
      tn = fold_node_fast(told, p, C->sgn[1]);
    else
      tn = fold_node_fast(told, p, C->sgn[0]);
    if (tn != told) T[c + g.sm] = tn;
    if (kReduce) d = fabs(tn - told);
  }
  if (kReduce) {
    const double v = wave_max_dt(d);
    if (tx == 0 && v > 0.0) max_dt_update(max_dt, v);
  }
}

// ---------------------------------------------------------------------------
// Split kernel.  A block owns one row of 64 target nodes and has kGroups
// waves; wave w evaluates units w, w + kGroups, ... (fold_unit) of all 64
// nodes with wave-uniform constants, and the partial minima meet in LDS.  A
// plane thus runs kGroups times as many waves as nodes/64, which is what the
// latency-bound node update needs to keep the SIMDs issuing.
// ---------------------------------------------------------------------------
// Units G0, G0 + kStride, ... (< 8) of one node, starting from the centre
// one-point value; unit numbers are compile-time constants.
template <typename R, int G0, int kStride>
__device__ __forceinline__ void units_of(R &tn, const Patch<R> &p, const FoldSign<R> &S) {
#pragma unroll
  for (int u = G0; u < 8; u += kStride) fold_unit(tn, p, S, u);
}

template <typename R, int G0, int kStride>
__device__ __forceinline__ R split_units(R told, const Patch<R> &p, const FoldFast<R> *C) {
  R tn;
  if (told < R(0)) {
    const FoldSign<R> &S = C->sgn[1];
    tn = fold_start(told, p.c, S);
    if (fold_gate_open(tn, p)) units_of<R, G0, kStride>(tn, p, S);
  } else {
    const FoldSign<R> &S = C->sgn[0];
    tn = fold_start(told, p.c, S);
    if (fold_gate_open(tn, p)) units_of<R, G0, kStride>(tn, p, S);
  }
  return tn;
}

template <typename R, int kGroups>
__device__ __forceinline__ R split_dispatch(int grp, R told, const Patch<R> &p,
                                            const FoldFast<R> *C) {
  if constexpr (kGroups == 2) {
    return grp == 0 ? split_units<R, 0, 2>(told, p, C) : split_units<R, 1, 2>(told, p, C);
  } else if constexpr (kGroups == 4) {
    switch (grp) {
      case 0: return split_units<R, 0, 4>(told, p, C);
      case 1: return split_units<R, 1, 4>(told, p, C);
      case 2: return split_units<R, 2, 4>(told, p, C);
      default: return split_units<R, 3, 4>(told, p, C);
    }
  } else {
    static_assert(kGroups == 8, "two, four or eight waves per node row");
    switch (grp) {
      case 0: return split_units<R, 0, 8>(told, p, C);
      case 1: return split_units<R, 1, 8>(told, p, C);
      case 2: return split_units<R, 2, 8>(told, p, C);
      case 3: return split_units<R, 3, 8>(told, p, C);
      case 4: return split_units<R, 4, 8>(told, p, C);
      case 5: return split_units<R, 5, 8>(told, p, C);
      case 6: return split_units<R, 6, 8>(told, p, C);
      default: return split_units<R, 7, 8>(told, p, C);
    }
  }
}

constexpr int kLockCounters = 256;

// Shape of a split-kernel block, by waves per row.  A wave row covers 2 node
// rows of a 32-node chunk (a wave: 64 nodes); a block holds split_rows() wave
// rows and stages their source rows plus a halo once.  Two wave rows for 2
// waves per row (4 waves per block), one otherwise.  Relative sweep times at
// 512^3 with 2 waves per row: 64-node chunks with one row per block 1.00, with
// two 0.98; 32-node chunks 0.93 (16-node chunks 0.94; 3 or 4 wave rows per
// block 1.14 and 1.06 against 2); with 4 and 8 waves per row (384^3, 256^3)
// 32-node chunks were 3 % faster than 64.  The blocks of a
// chunk column keep their XCD (linear block index mod 8), so with 32-node
// chunks each XCD owns every eighth chunk of every row: the work of a busy
// region spreads over more XCDs while each chunk's source rows stay in its
// XCD's L2 from plane to plane.
constexpr int split_rows(int groups) { return groups == 2 ? 2 : 1; }
constexpr int split_width(int) { return 32; }

// Exact locking (kLock).  A stamp is the sweep index sigma = 6*iteration +
// direction at which a node last changed (0 before any change), kept not per
// node but per line chunk: 64 consecutive nodes of one x-line (j, k fixed) or
// of one y-line (i, k fixed), in global coordinates, so the stamps need no
// moving when the field changes layout.  A node that changes writes sigma
// into its x-line chunk and its y-line chunk.  The rows of every layout are
// x-lines (reference and y-sweep layouts) or y-lines (x-sweep layout); the
// lanes of a row write one word of their own family and consecutive words of
// the other (x-line chunks are indexed (k, i/64, j), y-line chunks (k, j/64,
// i)).  A row of targets was last updated in this direction at sweep
// sigma - 6, after that sweep had finished the source plane; if no chunk its
// source patch touches has changed since (every stamp <= sigma - 6), its
// candidates are the ones it saw then, all >= its value, so the update would
// return the value unchanged and the row is skipped.  Same bits as without
// locking.
struct LineStamps {
  int *xs, *ys;   // x-line chunks ((k chx + i/64) (ny+2) + j), y-line chunks ((k chy + j/64) (nx+2) + i)
  int chx, chy;   // chunks per x-line, per y-line
  int hx, hy;     // nx + 2, ny + 2
};
__device__ __forceinline__ int *line_stamp(const LineStamps &L, bool yline, int i, int j, int k) {
  return yline ? L.ys + ((long long)k * L.chy + (j >> 6)) * L.hx + i
               : L.xs + ((long long)k * L.chx + (i >> 6)) * L.hy + j;
}
// Where one plane launch sits in global coordinates: the lane axis la (0 or 1:
// the rows are x-lines or y-lines), the row axis lb, the march axis m, the
// source and target plane indices.
struct LockGeom {
  LineStamps L;
  int la, lb, m;
  int psrc, ptgt;
  int sigma;
  unsigned long long *updates_computed;  // cell-updates computed (kLockCounters words)
};

// A block owns kRows wave rows of one kW-node chunk (blockIdx.x): a wave row
// is 64 / kW node rows of the chunk (lane l: node l % kW of row l / kW), and
// its kGroups waves are threadIdx.y = r * kGroups + grp.  Blocks go to the
// XCDs round-robin by their linear index, so a chunk stays on one XCD from
// plane to plane and its source rows are still in that XCD's L2 from the
// launch that wrote them.  (Rotating the chunk by the row group, so that each
// XCD gets every chunk, balanced the XCDs' work but was 3-4 % slower overall.)
// One block's rows (row group `by` of chunk column blockIdx.x); s and part
// are the block's LDS.
template <typename R, bool kReduce, int kGroups, bool kLock, int kRows, int kW>
__device__ __forceinline__ void split_rowgroup(R *__restrict__ T, const PlaneGeom &g,
                                               const FoldFast<R> *__restrict__ C,
                                               unsigned long long *__restrict__ max_dt,
                                               const LockGeom &lk, int by,
                                               R (&s)[kRows * (64 / kW) + 2][kW + 2],
                                               R (&part)[kRows][kGroups][64]) {
  constexpr int kRpw = 64 / kW;         // node rows per wave row
  constexpr int kNR = kRows * kRpw;     // node rows per block
  const int tx = threadIdx.x;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int grp = wy % kGroups, row = wy / kGroups;
  const int a0 = blockIdx.x * kW, b0 = 1 + by * kNR;
  int changed = 0;
  if (kLock && wy == 0 && tx < 2 * (kNR + 2)) {
    // the stamps of the 64-node line chunks the staged rows touch (columns
    // a0 .. a0 + kW + 1), loaded beside the staging loads
    const int r = tx >> 1, ca = (tx & 1) ? a0 + kW + 1 : a0, sb = b0 - 1 + r;
    if (ca <= g.na + 1 && sb <= g.nb + 1) {
      int c3[3];
      c3[lk.la] = ca;
      c3[lk.lb] = sb;
      c3[lk.m] = lk.psrc;
      changed = *line_stamp(lk.L, lk.la == 1, c3[0], c3[1], c3[2]) > lk.sigma - 6;
    }
  }
  // Locking: the block votes on its stamps before it loads anything else, so
  // a skipped block costs one small load and no store.  (Locking runs once
  // most rows are idle; loading the staged values beside the stamps instead
  // costs a busy block nothing but leaves an idle one waiting for them.)
  // Only the blocks that go on count their updates, at the end (counting the
  // skipped ones, one atomic per idle block, cost 6 % of a solve to the fixed
  // point).
  if (kLock && lk.sigma >= 6 && !__syncthreads_or(changed)) return;  // uniform over the block
  // Loads: the target value and the source rows (which go through registers
  // into LDS) are all in flight together.  Every lane loads, from an address
  // clamped into the padded grid, and the lanes outside the plane or the
  // stage drop what they read: a conditional load is a branch, and the
  // compiler waits for a load inside its branch when it moves the value's
  // first use (fabs) there, which made the two loads one after the other.
  const int la_ = tx % kW, lb_ = row * kRpw + tx / kW;  // the lane's node in the block
  const int a = a0 + 1 + la_, b = b0 + lb_;
  const bool active = a <= g.na && b <= g.nb;
  // (clamped: the target of an active lane, where only active lanes store)
  const long long c = g.base + (long long)min(a, g.na) * g.Sa + (long long)min(b, g.nb) * g.Sb;
  const R tload = T[c + g.sm];
  constexpr int kStage = (kNR + 2) * (kW + 2), kThreads = 64 * kGroups * kRows;
  constexpr int kPer = (kStage + kThreads - 1) / kThreads;
  R sv[kPer];
  bool sok[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = wy * 64 + tx + q * kThreads;
    const int la = idx % (kW + 2), lb = idx / (kW + 2);
    const int sa = a0 + la, sb = b0 - 1 + lb;
    sok[q] = idx < kStage && sa <= g.na + 1 && sb <= g.nb + 1;
    sv[q] = T[g.base + (long long)min(sa, g.na + 1) * g.Sa + (long long)min(sb, g.nb + 1) * g.Sb];
  }
  const R told = active ? tload : R(0);
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = wy * 64 + tx + q * kThreads;
    if (idx < kStage) s[idx / (kW + 2)][idx % (kW + 2)] = sok[q] ? fabs(sv[q]) : R(0);
  }
  __syncthreads();
  R tn = fabs(told);
  if (active) {
    // the patch around the lane's node: the offsets of +x^ and +y^ within the
    // staged rows are the same for every lane (uniform products)
    const R *c0 = &s[lb_ + 1][la_ + 1];
    const int ox = g.ub * (kW + 2) + g.ua, oy = g.vb * (kW + 2) + g.va;
    Patch<R> p;
    p.c = c0[0];
    p.px = c0[ox];
    p.py = c0[oy];
    p.mx = c0[-ox];
    p.my = c0[-oy];
    p.pxpy = c0[ox + oy];
    p.pxmy = c0[ox - oy];
    p.mxpy = c0[oy - ox];
    p.mxmy = c0[-ox - oy];
    tn = split_dispatch<R, kGroups>(grp, told, p, C);  // grp is wave-uniform
  }
  part[row][grp][tx] = tn;
  __syncthreads();
  if (kLock && wy == kGroups - 1 && tx == 0) {
    // computed updates, spread over kLockCounters addresses; issued by a wave
    // that has nothing left to do, so that no load waits behind the atomic
    const int nrows = g.nb - b0 + 1 < kNR ? g.nb - b0 + 1 : kNR;  // rows in the plane
    atomicAdd(&lk.updates_computed[(blockIdx.x + 7u * by) % kLockCounters],
              (unsigned long long)(g.na - a0 < kW ? g.na - a0 : kW) * nrows);
  }
  if (grp == 0) {
    R d = R(0);
    if (active) {
      R m = part[row][0][tx];
#pragma unroll
      for (int k = 1; k < kGroups; ++k) m = vmin(part[row][k][tx], m);
      const R res = ((told > R(0)) ? R(1) : R(-1)) * m;
      if (res != told) {
        T[c + g.sm] = res;
        if (kLock) {
          int c3[3];
          c3[lk.la] = a;
          c3[lk.lb] = b;
          c3[lk.m] = lk.ptgt;
          *line_stamp(lk.L, false, c3[0], c3[1], c3[2]) = lk.sigma;
          *line_stamp(lk.L, true, c3[0], c3[1], c3[2]) = lk.sigma;
        }
      }
      if (kReduce) d = fabs(res - told);
    }
    if (kReduce) {
      const double v = wave_max_dt(d);
      if (tx == 0 && v > 0.0) max_dt_update(max_dt, v);
    }
  }
}

template <typename R, bool kReduce, int kGroups, bool kLock, int kRows, int kW>
__global__ __launch_bounds__(64 * kGroups * kRows) void fold_split_kernel(
    R *__restrict__ T, PlaneGeom g, const FoldFast<R> *__restrict__ C,
    unsigned long long *__restrict__ max_dt, LockGeom lk) {
  constexpr int kNR = kRows * (64 / kW);
  __shared__ R s[kNR + 2][kW + 2];
  __shared__ R part[kRows][kGroups][64];
  // (Locked launches with fewer blocks, each taking 2, 4 or 8 row groups in a
  // loop, were tried for the mostly idle late iterations: the loop raised the
  // kernel to 99 VGPRs, and even the gain of 2 row groups in the last
  // iterations was below what the lost occupancy cost the busy ones.)
  split_rowgroup<R, kReduce, kGroups, kLock, kRows, kW>(T, g, C, max_dt, lk, blockIdx.y, s, part);
}

__global__ void fill_int_kernel(int *__restrict__ p, long long n, int v) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}

// Locking state of one solve (null stamps: no locking).
struct LockState {
  LineStamps L;                        // L.xs == nullptr: no locking
  int sigma;                           // sweep index of the launch
  unsigned long long *updates_computed;   // cell-updates computed by locked launches
};

template <typename R, int kGroups>
static void launch_split(hipStream_t st, R *T, const PlaneGeom &g, const FoldFast<R> *C,
                         unsigned long long *max_dt, const LockState &L, const LockGeom &lk) {
  constexpr int kR = split_rows(kGroups), kW = split_width(kGroups);
  constexpr int kNR = kR * (64 / kW);
  const dim3 block(64, kGroups * kR);
  const dim3 grid((g.na + kW - 1) / kW, (g.nb + kNR - 1) / kNR);
  if (L.L.xs) {
    if (max_dt)
      fold_split_kernel<R, true, kGroups, true, kR, kW><<<grid, block, 0, st>>>(T, g, C, max_dt,
                                                                              lk);
    else
      fold_split_kernel<R, false, kGroups, true, kR, kW><<<grid, block, 0, st>>>(T, g, C, nullptr,
                                                                               lk);
    return;
  }
  if (max_dt) {
    fold_split_kernel<R, true, kGroups, false, kR, kW><<<grid, block, 0, st>>>(T, g, C, max_dt,
                                                                             lk);
  } else {
    fold_split_kernel<R, false, kGroups, false, kR, kW><<<grid, block, 0, st>>>(T, g, C, nullptr,
                                                                              lk);
  }
}

// groups: 0 = one thread per node (fold_tile_kernel), 2/4/8 = waves per node row.
template <typename R>
static void tiled_direction(hipStream_t st, R *T, const Grid &G, int d,
                            const FoldFast<R> *C, unsigned long long *max_dt, int groups,
                            const LockState &L) {
  const Frame &fr = kFrames[d];
  const int n = G.n[fr.m];
  // lane and row axes of this layout's planes (as plane_geom picks them)
  const bool u_first = G.stride[fr.ux] < G.stride[fr.uy];
  LockGeom lk{L.L, u_first ? fr.ux : fr.uy, u_first ? fr.uy : fr.ux, fr.m, 0, 0, L.sigma,
              L.updates_computed};
  for (int q = 0; q < n - 1; ++q) {
    const int p = fr.s > 0 ? 1 + q : n - q;
    const PlaneGeom g = plane_geom(G, d, p);
    lk.psrc = p;
    lk.ptgt = p + fr.s;
    if (groups == 2) {
      launch_split<R, 2>(st, T, g, C, max_dt, L, lk);
    } else if (groups == 4) {
      launch_split<R, 4>(st, T, g, C, max_dt, L, lk);
    } else if (groups == 8) {
      launch_split<R, 8>(st, T, g, C, max_dt, L, lk);
    } else {
      const dim3 block(kTileA, kTileB);
      const dim3 grid((g.na + kTileA - 1) / kTileA, (g.nb + kTileB - 1) / kTileB);
      if (max_dt)
        fold_tile_kernel<R, true><<<grid, block, 0, st>>>(T, g, C, max_dt);
      else
        fold_tile_kernel<R, false><<<grid, block, 0, st>>>(T, g, C, nullptr);
    }
  }
}

template <typename R>
int fold_sweeps_tiled_engine(hipStream_t st, R *T, R *Y, const Grid &G,
                             const FoldFast<R> *d_consts, int max_iters, double tol, int groups,
                             unsigned long long *d_max, int *iters_out, double *max_dt_out,
                             int *stamps, unsigned long long *updates_computed,
                             int lock_from) {
  const Grid GX = G.with_order(kOrderX), GZ = G.with_order(kOrderZ), GY = G.with_order(kOrderY);
  const bool lock = stamps != nullptr && groups > 0;
  LineStamps L{nullptr, nullptr, 0, 0, G.n[0] + 2, G.n[1] + 2};
  if (lock) {
    L.chx = line_chunks(G.n[0]);
    L.chy = line_chunks(G.n[1]);
    L.xs = stamps;
    L.ys = stamps + line_stamp_count(G, false);
  }
  const LineStamps none{nullptr, nullptr, 0, 0, 0, 0};
  FieldMover<R> mv(st, T, Y, GZ);
  int it = 0;
  auto dir = [&](const Grid &Gd, int d, unsigned long long *md) {
    tiled_direction<R>(st, mv.field(), Gd, d, d_consts + d, md, groups,
                       LockState{lock && it >= lock_from ? L : none, 6 * it + d,
                                 updates_computed});
  };
  double last = -1.0;
  for (; it < max_iters;) {
    unsigned long long *md = tol > 0 ? d_max : nullptr;
    if (md) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    if (lock && it == lock_from) {
      // Locking starts: every chunk counts as changed at the last sweep before
      // (6 it - 1; 0 at the start, when sigma < 6 forbids skipping anyway), a
      // bound of the sweep of any change so far.
      fill_int_kernel<<<1024, 256, 0, st>>>(stamps, (long long)line_stamp_words(G),
                                            it > 0 ? 6 * it - 1 : 0);
    }
    // Two moves per iteration: x_down, x_up with x slowest; then the z- and
    // the y-sweeps with y slowest (a z = const plane there is rows of x-lines,
    // each contiguous, as the kernel needs), except in the last iteration of a
    // fixed-count solve, which runs them in the reference layout so that the
    // field ends in the caller's buffer after an even number of moves.
    mv.to(GX);
    dir(GX, FOLD3D_X_DOWN, md);
    dir(GX, FOLD3D_X_UP, md);
    const Grid &GP = tol <= 0 && it == max_iters - 1 ? GZ : GY;
    mv.to(GP);
    dir(GP, FOLD3D_K_UP, md);
    dir(GP, FOLD3D_K_DOWN, md);
    dir(GP, FOLD3D_Y_UP, md);
    dir(GP, FOLD3D_Y_DOWN, md);
    F3D_HIP(hipGetLastError());
    ++it;
    if (md) {
      F3D_HIP(max_dt_read(st, d_max, &last));
      if (last < tol) break;
    }
  }
  F3D_HIP(mv.finish(GZ));
  *iters_out = it;
  *max_dt_out = last;
  return FOLD3D_OK;
}
// The split kernel on the caller's field in the reference layout for all six
// directions: no second field and no layout moves (the x-sweeps' rows then
// run along y, n + 2 values apart, so their staging loads are not
// coalesced).  Same bits as the other engines: the kernel reads each plane
// through its strides.
template <typename R>
int fold_sweeps_inplace_engine(hipStream_t st, R *T, const Grid &G, const FoldFast<R> *d_consts,
                               int max_iters, double tol, int groups, unsigned long long *d_max,
                               int *iters_out, double *max_dt_out) {
  const LockState none{LineStamps{nullptr, nullptr, 0, 0, 0, 0}, 0, nullptr};
  int it = 0;
  double last = -1.0;
  for (; it < max_iters;) {
    unsigned long long *md = tol > 0 ? d_max : nullptr;
    if (md) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    for (int d = 0; d < 6; ++d) tiled_direction<R>(st, T, G, d, d_consts + d, md, groups, none);
    F3D_HIP(hipGetLastError());
    ++it;
    if (md) {
      F3D_HIP(max_dt_read(st, d_max, &last));
      if (last < tol) break;
    }
  }
  *iters_out = it;
  *max_dt_out = last;
  return FOLD3D_OK;
}
template int fold_sweeps_inplace_engine<double>(hipStream_t, double *, const Grid &,
                                                const FoldFast<double> *, int, double, int,
                                                unsigned long long *, int *, double *);
template int fold_sweeps_inplace_engine<float>(hipStream_t, float *, const Grid &,
                                               const FoldFast<float> *, int, double, int,
                                               unsigned long long *, int *, double *);

template int fold_sweeps_tiled_engine<double>(hipStream_t, double *, double *, const Grid &,
                                              const FoldFast<double> *, int, double, int,
                                              unsigned long long *, int *, double *, int *,
                                              unsigned long long *, int);
template int fold_sweeps_tiled_engine<float>(hipStream_t, float *, float *, const Grid &,
                                             const FoldFast<float> *, int, double, int,
                                             unsigned long long *, int *, double *, int *,
                                             unsigned long long *, int);

}  // namespace fold3d
