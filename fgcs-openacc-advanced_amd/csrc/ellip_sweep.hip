// ellip_sweep.hip — plane-step kernels and sweep driver of the tilted
// elliptic model (constant or per-node metric, constant or per-node speed),
// FP64 and FP32.  Same plane order and frames as the reference's fold-model
// driver (AN_ISO_1_OpenACC.c:1192-1237); convergence by max|dT| < tol.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "common.h"
#include "ellip_local.h"
#include "metric_prep.h"
#include "solver.h"

namespace fold3d {

constexpr int kLaneE = 64;
constexpr int kRowsE = 4;

template <typename R>
struct EllipArgs {
  EllipConsts<R> K;        // constant metric (unused with per-node fields)
  R speed;                 // constant speed (unused with a speed field)
  const R *speed_field;    // per node, padded layout, or null
  const R *theta, *phi, *ra, *rb;  // per node, or null
  int u[3];                // local frame axes x^, y^, z^ (global axis ids)
  R sg[3];                 // their signs
  R hl[3];                 // spacings along x^, y^, z^
  R gk;                    // constant metric: plane_gap of the direction for s = 1
};

template <typename R, bool kReduce, bool kSpeed, bool kMetric>
__global__ __launch_bounds__(kLaneE *kRowsE) void ellip_plane_kernel(
    R *__restrict__ T, PlaneGeom g, EllipArgs<R> A,
    unsigned long long *__restrict__ max_dt) {
  const int a = 1 + blockIdx.x * kLaneE + threadIdx.x;
  const int b = 1 + blockIdx.y * kRowsE + threadIdx.y;
  R d = R(0);
  if (a <= g.na && b <= g.nb) {
    const long long c = g.base + a * g.Sa + b * g.Sb;
    const long long t = c + g.sm;
    R p[9];
    p[0] = T[c];
#pragma unroll
    for (int k = 0; k < 8; ++k) p[1 + k] = T[c + kRing[k][0] * g.su + kRing[k][1] * g.sv];
    const R told = T[t];
    const R s = kSpeed ? R(1) / A.speed_field[t] : R(1) / A.speed;
    R tn;
    if (kMetric) {
      R N[6];
      tilted_inverse_metric<R>(A.theta[t], A.phi[t], A.ra[t], A.rb[t], N);
      const LatticeN<R> L = lattice_of<R>(N, A.u, A.sg, A.hl);
      tn = ellip_node_lattice<R>(told, p, s, L);
    } else {
      tn = ellip_node_update<R>(told, p, s, A.K);
    }
    if (tn != told) T[t] = tn;
    if (kReduce) d = told - tn;
  }
  if (kReduce) {
    const double v = wave_max_dt(d);
    if ((threadIdx.x & 63) == 0 && v > 0.0)
      max_dt_update(max_dt, v);
  }
}

template <typename R, bool kSpeed, bool kMetric>
static void ellip_launch(hipStream_t st, R *T, const PlaneGeom &g,
                         const EllipArgs<R> &A, unsigned long long *max_dt) {
  dim3 block(kLaneE, kRowsE);
  dim3 grid((g.na + kLaneE - 1) / kLaneE, (g.nb + kRowsE - 1) / kRowsE);
  if (max_dt)
    ellip_plane_kernel<R, true, kSpeed, kMetric><<<grid, block, 0, st>>>(T, g, A, max_dt);
  else
    ellip_plane_kernel<R, false, kSpeed, kMetric><<<grid, block, 0, st>>>(T, g, A, nullptr);
}

template <typename R>
static void ellip_dispatch(hipStream_t st, R *T, const PlaneGeom &g,
                           const EllipArgs<R> &A, unsigned long long *max_dt) {
  const bool sp = A.speed_field != nullptr, me = A.theta != nullptr;
  if (sp && me) ellip_launch<R, true, true>(st, T, g, A, max_dt);
  else if (sp) ellip_launch<R, true, false>(st, T, g, A, max_dt);
  else if (me) ellip_launch<R, false, true>(st, T, g, A, max_dt);
  else ellip_launch<R, false, false>(st, T, g, A, max_dt);
}

// Host-side constants of direction `dir` (double precision, then rounded to R).
template <typename R>
static EllipArgs<R> ellip_args(int dir, const fold3d_model &m, const double h[3]) {
  const Frame &fr = kFrames[dir];
  EllipArgs<R> A;
  memset(&A, 0, sizeof A);
  A.u[0] = fr.ux;
  A.u[1] = fr.uy;
  A.u[2] = fr.m;
  A.sg[0] = (R)fr.sx;
  A.sg[1] = (R)fr.sy;
  A.sg[2] = (R)fr.s;
  A.hl[0] = (R)h[fr.ux];
  A.hl[1] = (R)h[fr.uy];
  A.hl[2] = (R)h[fr.m];
  A.speed = (R)m.speed;
  A.speed_field = (const R *)m.speed_field;
  A.theta = (const R *)m.theta;
  A.phi = (const R *)m.phi;
  A.ra = (const R *)m.ra;
  A.rb = (const R *)m.rb;
  if (!m.theta) {
    double N[6];
    sym3_inverse(m.M, N);
    const int u[3] = {fr.ux, fr.uy, fr.m};
    const double sg[3] = {(double)fr.sx, (double)fr.sy, (double)fr.s};
    const double hl[3] = {h[fr.ux], h[fr.uy], h[fr.m]};
    EllipConsts<double> Kd;
    ellip_constants<double>(N, u, sg, hl, Kd);
    const double *src = &Kd.vert[0];
    R *dst = &A.K.vert[0];
    for (size_t i = 0; i < sizeof(Kd) / sizeof(double); ++i) dst[i] = (R)src[i];
    // plane_gap of the direction for unit slowness, as the device computes it
    // from the lattice products (the same guards and shrink factor)
    auto nl = [&](int i, int j) {
      const double Ng[3][3] = {{N[0], N[3], N[4]}, {N[3], N[1], N[5]}, {N[4], N[5], N[2]}};
      return sg[i] * sg[j] * Ng[u[i]][u[j]];
    };
    const double La = nl(0, 0) * hl[0] * hl[0], Lb = nl(1, 1) * hl[1] * hl[1];
    const double Lc = nl(2, 2) * hl[2] * hl[2], Le = nl(0, 1) * hl[0] * hl[1];
    const double Lfx = nl(0, 2) * hl[0] * hl[2], Lfy = nl(1, 2) * hl[1] * hl[2];
    const double det = La * Lb - Le * Le;
    const double q = det > 1e-6 * La * Lb
                         ? Lc - (Lb * Lfx * Lfx - 2.0 * Le * Lfx * Lfy + La * Lfy * Lfy) / det
                         : 0.0;
    A.gk = q > 1e-3 * Lc ? (R)((sizeof(R) == 8 ? 0.999 : 0.9) * std::sqrt(q)) : R(0);
  }
  return A;
}

void sym3_inverse(const double M[6], double N[6]) {
  const double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5];
  const double A = b * c - f * f, B = a * c - e * e, C = a * b - d * d;
  const double D = e * f - c * d, E = d * f - b * e, F = d * e - a * f;
  const double det = a * A + d * D + e * E;
  N[0] = A / det;
  N[1] = B / det;
  N[2] = C / det;
  N[3] = D / det;
  N[4] = E / det;
  N[5] = F / det;
}

template <typename R>
int ellip_sweeps_plane_engine(hipStream_t st, R *T, const Grid &G,
                              const fold3d_model &m, const double h[3],
                              int max_iters, double tol,
                              unsigned long long *d_max, int *iters_out,
                              double *max_dt_out) {
  EllipArgs<R> args[6];
  for (int d = 0; d < 6; ++d) args[d] = ellip_args<R>(d, m, h);
  int it = 0;
  double last = -1.0;
  for (; it < max_iters;) {
    if (tol > 0) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    for (int d = 0; d < 6; ++d) {
      const Frame &fr = kFrames[d];
      const int n = G.n[fr.m];
      unsigned long long *md = tol > 0 ? d_max : nullptr;
      if (fr.s > 0)
        for (int p = 1; p < n; ++p) ellip_dispatch<R>(st, T, plane_geom(G, d, p), args[d], md);
      else
        for (int p = n; p > 1; --p) ellip_dispatch<R>(st, T, plane_geom(G, d, p), args[d], md);
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

template int ellip_sweeps_plane_engine<double>(hipStream_t, double *, const Grid &,
                                               const fold3d_model &, const double *,
                                               int, double, unsigned long long *,
                                               int *, double *);
template int ellip_sweeps_plane_engine<float>(hipStream_t, float *, const Grid &,
                                              const fold3d_model &, const double *,
                                              int, double, unsigned long long *,
                                              int *, double *);

// ---------------------------------------------------------------------------
// Split engine for a constant metric.  As for the fold model: a block owns a
// row of 64 target nodes, stages the three source rows in LDS, and kGroups
// waves share the node's eight units (triangle k with spoke k and rim k); each
// wave also takes the nine one-point candidates so its pruning starts low.
// The partial minima meet in LDS.  Same bits as ellip_plane_kernel.
// ---------------------------------------------------------------------------
template <typename R, int U>
__device__ __forceinline__ void ellip_unit(R &best, const R p[9], R s2, const EllipConsts<R> &K,
                                           R gap) {
  constexpr int k1 = 1 + U, k2 = 1 + (U + 1) % 8;
  if (!may_win<R, true>(rmin(p[0], rmin(p[k1], p[k2])), gap, best)) return;
  // the unit's twelve constants in scalar registers at once (one wait, not
  // one per simplex: the compiler would sink each load into its branch)
  asm volatile("" ::"s"(K.edge[U][0]), "s"(K.edge[U][1]), "s"(K.edge[U][2]),
               "s"(K.edge[8 + U][0]), "s"(K.edge[8 + U][1]), "s"(K.edge[8 + U][2]),
               "s"(K.tri[U][0]), "s"(K.tri[U][1]), "s"(K.tri[U][2]), "s"(K.tri[U][3]),
               "s"(K.tri[U][4]), "s"(K.tri[U][5]));
  ellip_edge<R, true>(best, p[0], p[k1], s2, K.edge[U], gap);
  ellip_edge<R, true>(best, p[k1], p[k2], s2, K.edge[8 + U], gap);
  ellip_tri<R, true>(best, p[0], p[k1], p[k2], s2, K.tri[U], gap);
}

template <typename R, int G0, int kStride>
__device__ __forceinline__ void ellip_units(R &best, const R p[9], R s2, const EllipConsts<R> &K,
                                            R gap) {
  if constexpr (G0 < 8) {
    ellip_unit<R, G0>(best, p, s2, K, gap);
    ellip_units<R, G0 + kStride, kStride>(best, p, s2, K, gap);
  }
}

template <typename R, int kGroups>
__device__ __forceinline__ void ellip_dispatch_units(int grp, R &best, const R p[9], R s2,
                                                     const EllipConsts<R> &K, R gap) {
  if constexpr (kGroups == 1) {
    ellip_units<R, 0, 1>(best, p, s2, K, gap);
  } else if constexpr (kGroups == 2) {
    if (grp == 0) ellip_units<R, 0, 2>(best, p, s2, K, gap);
    else ellip_units<R, 1, 2>(best, p, s2, K, gap);
  } else {
    static_assert(kGroups == 4, "one, two or four waves per node row");
    switch (grp) {
      case 0: ellip_units<R, 0, 4>(best, p, s2, K, gap); break;
      case 1: ellip_units<R, 1, 4>(best, p, s2, K, gap); break;
      case 2: ellip_units<R, 2, 4>(best, p, s2, K, gap); break;
      default: ellip_units<R, 3, 4>(best, p, s2, K, gap); break;
    }
  }
}

// x = 1 / x in place (the slowness field from the speed field; the same
// correctly rounded division the plane kernel does with kSpeed == 1).
template <typename R>
__global__ void reciprocal_kernel(R *__restrict__ p, long long n) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = R(1) / p[k];
}

// kSpeed: 0 the model's constant speed, 1 a speed field V (slowness 1 / V per
// node), 2 a slowness field (already 1 / V, made once per solve).
template <typename R, bool kReduce, int kSpeed, int kGroups>
__global__ __launch_bounds__(64 * kGroups) void ellip_split_kernel(
    R *__restrict__ T, const R *__restrict__ V, PlaneGeom g, EllipArgs<R> A,
    unsigned long long *__restrict__ max_dt) {
  __shared__ R s[3][64 + 2];
  __shared__ R part[kGroups][64];
  const int tx = threadIdx.x;
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int a0 = blockIdx.x * 64, b = 1 + blockIdx.y;
  // all loads of the step in flight together (RowStage); a lane outside the
  // plane reads the block's first target node instead, and drops it
  const int a = a0 + 1 + tx;
  const bool active = a <= g.na;
  const long long c = g.base + (long long)a * g.Sa + (long long)b * g.Sb;
  const long long t = c + g.sm;
  const long long tl = active ? t : t + (long long)(a0 + 1 - a) * g.Sa;
  const R tload = T[tl];
  const R vload = kSpeed ? V[tl] : R(1);
  RowStage<R, 64 * kGroups> rs;
  rs.load(T, g.base, g.Sa, g.Sb, a0, b, g.na, grp * 64 + tx);
  rs.store(s, grp * 64 + tx);
  const R told = active ? tload : R(0);
  const R sl = kSpeed == 2   ? (active ? vload : R(1))
               : kSpeed == 1 ? (active ? R(1) / vload : R(1))
                             : R(1) / A.speed;
  __syncthreads();
  R best = told;
  if (active) {
    // the patch: the offsets of +x^ and +y^ within the staged rows are the
    // same for every lane (uniform products)
    const R *c0 = &s[1][tx + 1];
    const int ox = g.ub * 66 + g.ua, oy = g.vb * 66 + g.va;
    R p[9];
    p[0] = c0[0];
#pragma unroll
    for (int k = 0; k < 8; ++k) p[1 + k] = c0[kRing[k][0] * ox + kRing[k][1] * oy];
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (reached(p[k])) best = rmin(ellip_vertex(p[k], sl, A.K.vert[k]), best);
    ellip_dispatch_units<R, kGroups>(grp, best, p, sl * sl, A.K, sl * A.gk);
  }
  part[grp][tx] = best;
  __syncthreads();
  if (grp == 0) {
    R d = R(0);
    if (active) {
      R m = part[0][tx];
#pragma unroll
      for (int k = 1; k < kGroups; ++k) m = rmin(part[k][tx], m);
      if (m != told) T[t] = m;
      if (kReduce) d = told - m;
    }
    if (kReduce) {
      const double v = wave_max_dt(d);
      if (tx == 0 && v > 0.0) max_dt_update(max_dt, v);
    }
  }
}

template <typename R, int kSpeed, int kGroups>
static void ellip_split_launch(hipStream_t st, R *T, const R *V, const PlaneGeom &g,
                               const EllipArgs<R> &A, unsigned long long *max_dt) {
  const dim3 block(64, kGroups);
  const dim3 grid((g.na + 63) / 64, g.nb);
  if (max_dt)
    ellip_split_kernel<R, true, kSpeed, kGroups><<<grid, block, 0, st>>>(T, V, g, A, max_dt);
  else
    ellip_split_kernel<R, false, kSpeed, kGroups><<<grid, block, 0, st>>>(T, V, g, A, nullptr);
}

template <typename R>
static void ellip_split_direction(hipStream_t st, R *T, const R *V, const Grid &G, int d,
                                  const EllipArgs<R> &A, unsigned long long *max_dt,
                                  int groups, bool slowness = false) {
  const Frame &fr = kFrames[d];
  const int n = G.n[fr.m];
  for (int q = 0; q < n - 1; ++q) {
    const int p = fr.s > 0 ? 1 + q : n - q;
    const PlaneGeom g = plane_geom(G, d, p);
    if (V && slowness) {
      if (groups == 1) ellip_split_launch<R, 2, 1>(st, T, V, g, A, max_dt);
      else if (groups == 2) ellip_split_launch<R, 2, 2>(st, T, V, g, A, max_dt);
      else ellip_split_launch<R, 2, 4>(st, T, V, g, A, max_dt);
    } else if (V) {
      if (groups == 1) ellip_split_launch<R, 1, 1>(st, T, V, g, A, max_dt);
      else if (groups == 2) ellip_split_launch<R, 1, 2>(st, T, V, g, A, max_dt);
      else ellip_split_launch<R, 1, 4>(st, T, V, g, A, max_dt);
    } else {
      if (groups == 1) ellip_split_launch<R, 0, 1>(st, T, V, g, A, max_dt);
      else if (groups == 2) ellip_split_launch<R, 0, 2>(st, T, V, g, A, max_dt);
      else ellip_split_launch<R, 0, 4>(st, T, V, g, A, max_dt);
    }
  }
}

template <typename R>
int ellip_sweeps_split_engine(hipStream_t st, R *T, R *Y, R *VX, R *VY, const Grid &G,
                              const fold3d_model &m, const double h[3], int max_iters,
                              double tol, int groups, unsigned long long *d_max,
                              int *iters_out, double *max_dt_out) {
  EllipArgs<R> args[6];
  for (int d = 0; d < 6; ++d) args[d] = ellip_args<R>(d, m, h);
  const Grid GX = G.with_order(kOrderX), GZ = G.with_order(kOrderZ), GY = G.with_order(kOrderY);
  const R *V = (const R *)m.speed_field;
  if (V) {  // the slowness 1 / V in the x- and y-sweep layouts, once per solve
    permute_layout<R>(st, V, GZ, VX, GX);
    permute_layout<R>(st, V, GZ, VY, GY);
    reciprocal_kernel<R><<<2048, 256, 0, st>>>(VX, G.count);
    reciprocal_kernel<R><<<2048, 256, 0, st>>>(VY, G.count);
  }
  FieldMover<R> mv(st, T, Y, GZ);
  int it = 0;
  double last = -1.0;
  for (; it < max_iters;) {
    unsigned long long *md = tol > 0 ? d_max : nullptr;
    if (md) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    // two moves per iteration, as in the fold engines: the x-sweeps with x
    // slowest, the z- and y-sweeps with y slowest (the reference layout in the
    // last iteration of a fixed-count solve)
    mv.to(GX);
    ellip_split_direction<R>(st, mv.field(), V ? VX : nullptr, GX, FOLD3D_X_DOWN,
                             args[FOLD3D_X_DOWN], md, groups, true);
    ellip_split_direction<R>(st, mv.field(), V ? VX : nullptr, GX, FOLD3D_X_UP,
                             args[FOLD3D_X_UP], md, groups, true);
    const bool in_z = tol <= 0 && it == max_iters - 1;
    const Grid &GP = in_z ? GZ : GY;
    // (the reference layout's pass reads the caller's speed field: V, not 1 / V)
    const R *VP = V ? (in_z ? V : VY) : nullptr;
    const bool slow = !in_z;
    mv.to(GP);
    ellip_split_direction<R>(st, mv.field(), VP, GP, FOLD3D_K_UP, args[FOLD3D_K_UP], md, groups,
                             slow);
    ellip_split_direction<R>(st, mv.field(), VP, GP, FOLD3D_K_DOWN, args[FOLD3D_K_DOWN], md,
                             groups, slow);
    ellip_split_direction<R>(st, mv.field(), VP, GP, FOLD3D_Y_UP, args[FOLD3D_Y_UP], md, groups,
                             slow);
    ellip_split_direction<R>(st, mv.field(), VP, GP, FOLD3D_Y_DOWN, args[FOLD3D_Y_DOWN], md,
                             groups, slow);
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

// ---------------------------------------------------------------------------
// Metric engine: per-node metric (tilt and axis-ratio fields).  The slowness
// and the six products P_ab = (N_ab h_a) h_b of N = M^-1 with the spacings
// (lattice_products) are computed once per solve straight into seven fields
// of the x-sweep layout and seven of the y-sweep layout (metric_prep_kernel,
// metric_prep.h); the plane steps are the split kernel's (LDS rows, units
// over kGroups waves) with the node's lattice products read from those
// fields.  Same bits as ellip_plane_kernel with kMetric.
// ---------------------------------------------------------------------------
// Where each lattice product of one direction comes from: L.X = sg * P_q.
template <typename R>
struct LatticeMap {
  const R *src[6];  // P_q field in the current layout, for A, B, C, E, Fx, Fy
  R sg[6];
};

template <typename R>
static LatticeMap<R> lattice_map(int d, const R *fields, long long count) {
  const Frame &fr = kFrames[d];
  const int u[3] = {fr.ux, fr.uy, fr.m};
  const int sgn[3] = {fr.sx, fr.sy, fr.s};
  const int pairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  LatticeMap<R> M;
  for (int e = 0; e < 6; ++e) {
    const int i = pairs[e][0], j = pairs[e][1];
    const int gi = u[i], gj = u[j];
    int q;  // symmetric storage (xx, yy, zz, xy, xz, yz)
    if (gi == gj) q = gi;
    else q = (gi + gj == 1) ? 3 : (gi + gj == 2) ? 4 : 5;
    M.src[e] = fields + (1 + q) * count;
    M.sg[e] = R(sgn[i] * sgn[j]);
  }
  return M;
}

template <typename R, int G0, int kStride>
__device__ __forceinline__ void lattice_units(R &best, const R p[9], R s2, const LatticeN<R> &L,
                                              R gap) {
  if constexpr (G0 < 8) {
    ellip_unit_lattice<R, G0, true>(best, p, s2, L, gap);
    lattice_units<R, G0 + kStride, kStride>(best, p, s2, L, gap);
  }
}

template <typename R, int kGroups>
__device__ __forceinline__ void lattice_dispatch_units(int grp, R &best, const R p[9], R s2,
                                                       const LatticeN<R> &L, R gap) {
  if constexpr (kGroups == 1) {
    lattice_units<R, 0, 1>(best, p, s2, L, gap);
  } else if constexpr (kGroups == 2) {
    if (grp == 0) lattice_units<R, 0, 2>(best, p, s2, L, gap);
    else lattice_units<R, 1, 2>(best, p, s2, L, gap);
  } else {
    static_assert(kGroups == 4, "one, two or four waves per node row");
    switch (grp) {
      case 0: lattice_units<R, 0, 4>(best, p, s2, L, gap); break;
      case 1: lattice_units<R, 1, 4>(best, p, s2, L, gap); break;
      case 2: lattice_units<R, 2, 4>(best, p, s2, L, gap); break;
      default: lattice_units<R, 3, 4>(best, p, s2, L, gap); break;
    }
  }
}

template <typename R, bool kReduce, int kGroups>
__global__ __launch_bounds__(64 * kGroups) void ellip_metric_kernel(
    R *__restrict__ T, const R *__restrict__ S, LatticeMap<R> Lm, PlaneGeom g,
    unsigned long long *__restrict__ max_dt) {
  __shared__ R s[3][64 + 2];
  __shared__ R part[kGroups][64];
  const int tx = threadIdx.x;
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int a0 = blockIdx.x * 64, b = 1 + blockIdx.y;
  // all loads of the step in flight together (RowStage); a lane outside the
  // plane reads the block's first target node instead, and drops it
  const int a = a0 + 1 + tx;
  const bool active = a <= g.na;
  const long long c = g.base + (long long)a * g.Sa + (long long)b * g.Sb;
  const long long t = c + g.sm;
  const long long tl = active ? t : t + (long long)(a0 + 1 - a) * g.Sa;
  const R tload = T[tl], sload = S[tl];
  R Lv[6];
#pragma unroll
  for (int e = 0; e < 6; ++e) Lv[e] = Lm.src[e][tl];
  RowStage<R, 64 * kGroups> rs;
  rs.load(T, g.base, g.Sa, g.Sb, a0, b, g.na, grp * 64 + tx);
  rs.store(s, grp * 64 + tx);
  const R told = active ? tload : R(0), sl = active ? sload : R(1);
  LatticeN<R> L;
  L.A = Lm.sg[0] * Lv[0];
  L.B = Lm.sg[1] * Lv[1];
  L.C = Lm.sg[2] * Lv[2];
  L.E = Lm.sg[3] * Lv[3];
  L.Fx = Lm.sg[4] * Lv[4];
  L.Fy = Lm.sg[5] * Lv[5];
  __syncthreads();
  R best = told;
  if (active) {
    // the patch: the offsets of +x^ and +y^ within the staged rows are the
    // same for every lane (uniform products)
    const R *c0 = &s[1][tx + 1];
    const int ox = g.ub * 66 + g.ua, oy = g.vb * 66 + g.va;
    R p[9];
    p[0] = c0[0];
#pragma unroll
    for (int k = 0; k < 8; ++k) p[1 + k] = c0[kRing[k][0] * ox + kRing[k][1] * oy];
    ellip_vertices_lattice<R>(best, p, sl, L);
    lattice_dispatch_units<R, kGroups>(grp, best, p, sl * sl, L, plane_gap(L, sl));
  }
  part[grp][tx] = best;
  __syncthreads();
  if (grp == 0) {
    R d = R(0);
    if (active) {
      R m = part[0][tx];
#pragma unroll
      for (int k = 1; k < kGroups; ++k) m = rmin(part[k][tx], m);
      if (m != told) T[t] = m;
      if (kReduce) d = told - m;
    }
    if (kReduce) {
      const double v = wave_max_dt(d);
      if (tx == 0 && v > 0.0) max_dt_update(max_dt, v);
    }
  }
}

template <typename R, int kGroups>
static void ellip_metric_launch(hipStream_t st, R *T, const R *S, const LatticeMap<R> &Lm,
                                const PlaneGeom &g, unsigned long long *max_dt) {
  const dim3 block(64, kGroups);
  const dim3 grid((g.na + 63) / 64, g.nb);
  if (max_dt)
    ellip_metric_kernel<R, true, kGroups><<<grid, block, 0, st>>>(T, S, Lm, g, max_dt);
  else
    ellip_metric_kernel<R, false, kGroups><<<grid, block, 0, st>>>(T, S, Lm, g, nullptr);
}

template <typename R>
static void ellip_metric_direction(hipStream_t st, R *T, const R *fields, const Grid &G, int d,
                                   const double h[3], unsigned long long *max_dt, int groups) {
  const Frame &fr = kFrames[d];
  const int n = G.n[fr.m];
  const LatticeMap<R> Lm = lattice_map<R>(d, fields, G.count);
  for (int q = 0; q < n - 1; ++q) {
    const int p = fr.s > 0 ? 1 + q : n - q;
    const PlaneGeom g = plane_geom(G, d, p);
    if (groups == 1) ellip_metric_launch<R, 1>(st, T, fields, Lm, g, max_dt);
    else if (groups == 2) ellip_metric_launch<R, 2>(st, T, fields, Lm, g, max_dt);
    else ellip_metric_launch<R, 4>(st, T, fields, Lm, g, max_dt);
  }
}

size_t ellip_metric_scratch_fields() { return 1 + 2 * kMetricFields; }

template <typename R>
int ellip_sweeps_metric_engine(hipStream_t st, R *T, R *scratch, const Grid &G,
                               const fold3d_model &m, const double h[3], int max_iters,
                               double tol, int groups, unsigned long long *d_max,
                               int *iters_out, double *max_dt_out) {
  const long long count = G.count;
  R *Y = scratch;                        // T in the x- or y-sweep layout
  R *MX = scratch + count;               // metric fields, x-sweep layout
  R *MY = MX + kMetricFields * count;    // metric fields, y-sweep layout
  const Grid GX = G.with_order(kOrderX), GZ = G.with_order(kOrderZ), GY = G.with_order(kOrderY);
  double N[6];
  sym3_inverse(m.M, N);
  launch_metric_prep<R>(st, count, G.n, GZ.stride, GX.stride, GY.stride,
                        (const R *)m.speed_field, (R)m.speed, (const R *)m.theta,
                        (const R *)m.phi, (const R *)m.ra, (const R *)m.rb, N, h, MX, MY);
  F3D_HIP(hipGetLastError());
  FieldMover<R> mv(st, T, Y, GZ);
  int it = 0;
  double last = -1.0;
  for (; it < max_iters;) {
    unsigned long long *md = tol > 0 ? d_max : nullptr;
    if (md) max_dt_reset_kernel<<<1, kMaxSlots, 0, st>>>(d_max);
    // two moves per iteration, as in the fold engines: the x-sweeps with x
    // slowest, the z- and y-sweeps with y slowest
    mv.to(GX);
    ellip_metric_direction<R>(st, mv.field(), MX, GX, FOLD3D_X_DOWN, h, md, groups);
    ellip_metric_direction<R>(st, mv.field(), MX, GX, FOLD3D_X_UP, h, md, groups);
    mv.to(GY);
    ellip_metric_direction<R>(st, mv.field(), MY, GY, FOLD3D_K_UP, h, md, groups);
    ellip_metric_direction<R>(st, mv.field(), MY, GY, FOLD3D_K_DOWN, h, md, groups);
    ellip_metric_direction<R>(st, mv.field(), MY, GY, FOLD3D_Y_UP, h, md, groups);
    ellip_metric_direction<R>(st, mv.field(), MY, GY, FOLD3D_Y_DOWN, h, md, groups);
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

template int ellip_sweeps_metric_engine<double>(hipStream_t, double *, double *, const Grid &,
                                                const fold3d_model &, const double *, int,
                                                double, int, unsigned long long *, int *,
                                                double *);
template int ellip_sweeps_metric_engine<float>(hipStream_t, float *, float *, const Grid &,
                                               const fold3d_model &, const double *, int, double,
                                               int, unsigned long long *, int *, double *);

template int ellip_sweeps_split_engine<double>(hipStream_t, double *, double *, double *,
                                               double *, const Grid &, const fold3d_model &,
                                               const double *, int, double, int,
                                               unsigned long long *, int *, double *);
template int ellip_sweeps_split_engine<float>(hipStream_t, float *, float *, float *, float *,
                                              const Grid &, const fold3d_model &,
                                              const double *, int, double, int,
                                              unsigned long long *, int *, double *);

}  // namespace fold3d
