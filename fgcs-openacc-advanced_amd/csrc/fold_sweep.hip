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
    double v = (double)d;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
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
    double v = (double)d;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
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
    double v = (double)d;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
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

// Locking (kLock): stamp[x] is the sweep index sigma = 6*iteration + direction
// at which node x last changed (0 before any change).  A row of targets was
// last updated in this direction at sweep sigma - 6, after that sweep had
// finished the source plane; if no node of its three source rows has changed
// since (every stamp <= sigma - 6), its candidates are the ones it saw then,
// all >= its value, so the update would return the value unchanged and the
// row is skipped.  Same bits as without locking.
template <typename R, bool kReduce, int kGroups, bool kLock>
__global__ __launch_bounds__(64 * kGroups) void fold_split_kernel(
    R *__restrict__ T, PlaneGeom g, const FoldFast<R> *__restrict__ C,
    unsigned long long *__restrict__ max_dt, int *__restrict__ stamp, int sigma,
    unsigned long long *__restrict__ updates_done) {
  __shared__ R s[3][64 + 2];
  __shared__ R part[kGroups][64];
  const int tx = threadIdx.x;
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int a0 = blockIdx.x * 64, b = 1 + blockIdx.y;
  int changed = 0;
  for (int idx = grp * 64 + tx; idx < 3 * 66; idx += 64 * kGroups) {
    const int la = idx % 66, lb = idx / 66;
    const int sa = a0 + la;
    if (sa <= g.na + 1) {
      const long long x = g.base + (long long)sa * g.Sa + (long long)(b - 1 + lb) * g.Sb;
      s[lb][la] = fabs(T[x]);
      if (kLock) changed |= stamp[x] > sigma - 6;
    }
  }
  if (kLock) {
    if (sigma >= 6 && !__syncthreads_or(changed)) return;  // uniform over the block
    if (tx == 0 && grp == 0)  // spread over kLockCounters addresses
      atomicAdd(&updates_done[(blockIdx.x + 7u * blockIdx.y) % kLockCounters],
                (unsigned long long)(g.na - a0 < 64 ? g.na - a0 : 64));
  }
  const int a = a0 + 1 + tx;
  const bool active = a <= g.na;
  const long long c = g.base + (long long)a * g.Sa + (long long)b * g.Sb;
  const R told = active ? T[c + g.sm] : R(0);
  __syncthreads();
  R tn = fabs(told);
  if (active) {
    const int ca = tx + 1;
    const int ua = g.ua, ub = g.ub, va = g.va, vb = g.vb;
    Patch<R> p;
    p.c = s[1][ca];
    p.px = s[1 + ub][ca + ua];
    p.py = s[1 + vb][ca + va];
    p.mx = s[1 - ub][ca - ua];
    p.my = s[1 - vb][ca - va];
    p.pxpy = s[1 + ub + vb][ca + ua + va];
    p.pxmy = s[1 + ub - vb][ca + ua - va];
    p.mxpy = s[1 - ub + vb][ca - ua + va];
    p.mxmy = s[1 - ub - vb][ca - ua - va];
    tn = split_dispatch<R, kGroups>(grp, told, p, C);  // grp is wave-uniform
  }
  part[grp][tx] = tn;
  __syncthreads();
  if (grp == 0) {
    R d = R(0);
    if (active) {
      R m = part[0][tx];
#pragma unroll
      for (int k = 1; k < kGroups; ++k) m = rmin(part[k][tx], m);
      const R res = ((told > R(0)) ? R(1) : R(-1)) * m;
      if (res != told) {
        T[c + g.sm] = res;
        if (kLock) stamp[c + g.sm] = sigma;
      }
      if (kReduce) d = fabs(res - told);
    }
    if (kReduce) {
      double v = (double)d;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
      if (tx == 0 && v > 0.0) max_dt_update(max_dt, v);
    }
  }
}

// Locking state of one solve (null stamps: no locking).
struct LockState {
  int *stamp;                          // current layout
  int sigma;                           // sweep index of the launch
  unsigned long long *updates_done;    // cell-updates computed (device counter)
};

template <typename R, int kGroups>
static void launch_split(hipStream_t st, R *T, const PlaneGeom &g, const FoldFast<R> *C,
                         unsigned long long *max_dt, const LockState &L) {
  const dim3 block(64, kGroups);
  const dim3 grid((g.na + 63) / 64, g.nb);
  if (L.stamp) {
    if (max_dt)
      fold_split_kernel<R, true, kGroups, true><<<grid, block, 0, st>>>(T, g, C, max_dt, L.stamp,
                                                                         L.sigma, L.updates_done);
    else
      fold_split_kernel<R, false, kGroups, true><<<grid, block, 0, st>>>(T, g, C, nullptr, L.stamp,
                                                                          L.sigma, L.updates_done);
  } else if (max_dt) {
    fold_split_kernel<R, true, kGroups, false><<<grid, block, 0, st>>>(T, g, C, max_dt, nullptr, 0,
                                                                        nullptr);
  } else {
    fold_split_kernel<R, false, kGroups, false><<<grid, block, 0, st>>>(T, g, C, nullptr, nullptr,
                                                                         0, nullptr);
  }
}

// groups: 0 = one thread per node (fold_tile_kernel), 2/4/8 = waves per node row.
template <typename R>
static void tiled_direction(hipStream_t st, R *T, const Grid &G, int d,
                            const FoldFast<R> *C, unsigned long long *max_dt, int groups,
                            const LockState &L) {
  const Frame &fr = kFrames[d];
  const int n = G.n[fr.m];
  for (int q = 0; q < n - 1; ++q) {
    const int p = fr.s > 0 ? 1 + q : n - q;
    const PlaneGeom g = plane_geom(G, d, p);
    if (groups == 2) {
      launch_split<R, 2>(st, T, g, C, max_dt, L);
    } else if (groups == 4) {
      launch_split<R, 4>(st, T, g, C, max_dt, L);
    } else if (groups == 8) {
      launch_split<R, 8>(st, T, g, C, max_dt, L);
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
                             int *stamp, int *stampY, unsigned long long *updates_done) {
  const Grid GX = G.with_order(kOrderX), GZ = G.with_order(kOrderZ), GY = G.with_order(kOrderY);
  const bool lock = stamp != nullptr && groups > 0;
  if (lock) F3D_HIP(hipMemsetAsync(stamp, 0, sizeof(int) * G.count, st));
  FieldMover<R> mv(st, T, Y, GZ, lock ? stamp : nullptr, lock ? stampY : nullptr);
  const bool extra = FieldMover<R>::extra_move_first(max_iters, tol);
  auto dir = [&](const Grid &Gd, int it, int d, unsigned long long *md) {
    tiled_direction<R>(st, mv.field(), Gd, d, d_consts + d, md, groups,
                       LockState{lock ? mv.stamps() : nullptr, 6 * it + d, updates_done});
  };
  int it = 0;
  double last = -1.0;
  for (; it < max_iters;) {
    unsigned long long *md = tol > 0 ? d_max : nullptr;
    if (md) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    mv.to(GX);  // x_down, x_up with x slowest
    dir(GX, it, FOLD3D_X_DOWN, md);
    dir(GX, it, FOLD3D_X_UP, md);
    mv.to(GZ);  // k_up, k_down in the reference layout
    dir(GZ, it, FOLD3D_K_UP, md);
    dir(GZ, it, FOLD3D_K_DOWN, md);
    mv.to(GY);  // y_up, y_down with y slowest
    dir(GY, it, FOLD3D_Y_UP, md);
    dir(GY, it, FOLD3D_Y_DOWN, md);
    if (it == 0 && extra) mv.to(GZ);
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
  const LockState none{nullptr, 0, nullptr};
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
                                              unsigned long long *, int *, double *, int *, int *,
                                              unsigned long long *);
template int fold_sweeps_tiled_engine<float>(hipStream_t, float *, float *, const Grid &,
                                             const FoldFast<float> *, int, double, int,
                                             unsigned long long *, int *, double *, int *, int *,
                                             unsigned long long *);

}  // namespace fold3d
