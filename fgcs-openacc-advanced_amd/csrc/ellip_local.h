// ellip_local.h — device local solver of the tilted elliptic model
//     sqrt(grad T^T M grad T) = s(x) = 1/v(x),   M symmetric positive definite,
// on the same 3x3 upwind patch and sweep frames as the reference's fold model
// (AN_ISO_1_OpenACC.c:195-326).  The reference has no such model.
//
// With N = M^-1, a straight segment d costs s*|d|_N.  The new value of target X
// is min(T_old, min over the patch's simplices S of min_{Y in S} T_lin(Y) +
// s|X - Y|_N).  The simplices are the 9 nodes, 16 edges (8 spokes from the
// centre, 8 rim edges) and 8 triangles fanned around the centre.  For a
// simplex with base Y0, edge vectors E and D = X - Y0 the constants
//     P = (E^T N E)^-1,  l0 = P E^T N D,  kappa = D^T N D - (E^T N D).l0
// depend only on the metric and the spacings; per node only the value
// differences delta_i = T_i - T_0 enter:
//     mu = P delta, r = s^2 - delta.mu, root = sqrt(kappa r),
//     lambda = l0 - (root/r) mu in the simplex  =>  candidate T_0 + delta.l0 + root.
// A candidate of S is never below min over S's vertices, so a simplex whose
// smallest vertex is not below the running minimum is skipped: pruning that
// cannot change the result.  The split, metric and batched kernels prune
// harder: S lies in the upwind plane, so a candidate is also at least that
// minimum plus s times the N-distance from X to the plane (plane_gap below).
// The one-thread-per-node plane kernel keeps the plain test, and the tests
// hold the two to the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include "fold_local.h"  // rmin

namespace fold3d {

// In-plane ring order around the centre, in units of (dxx, dyy):
// +x, +x+y, +y, -x+y, -x, -x-y, -y, +x-y.
constexpr int kRing[8][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1},
                             {-1, 0}, {-1, -1}, {0, -1}, {1, -1}};

template <typename R>
struct EllipConsts {
  R vert[9];      // |X - Y|_N, centre then ring
  R edge[16][3];  // P, l0, kappa; 0..7 spokes centre->ring k, 8..15 rims ring k -> k+1
  R tri[8][6];    // P11, P12, P22, l1, l2, kappa; centre, ring k, ring k+1
};

// Unreached nodes carry MAXT (1e16); a simplex touching one is not used.
template <typename R>
__device__ __forceinline__ bool reached(R t) { return t < R(1e15); }

// The pruning test: can a simplex with smallest vertex `lo` improve on `best`?
// kGap: the bound adds `gap` (a lower bound of s |X - Y|_N over the upwind
// plane, already shrunk by plane_gap) and allows 16 ulps of `best` for the
// candidate's own rounding.  Without kGap: the plain vertex test.
template <typename R, bool kGap>
__device__ __forceinline__ bool may_win(R lo, R gap, R best) {
  if constexpr (kGap) return lo + gap < best * prune_slack<R>();
  return lo < best;
}

// ---------------------------------------------------------------------------
// Candidate arithmetic.  The elliptic kernels evaluate a simplex's candidate
// with fused multiply-adds at fixed places (ellip_edge_lat, ellip_tri_lat
// below; all of them in FP64, the per-node-metric ones in FP32), and the
// elliptic oracle restates the same operations, so the FP64 constant-metric
// kernels and the oracle agree bit for bit.  The square root is the correctly
// rounded one in FP64 and the hardware's in FP32.  The
// per-node-metric (lattice) path also forms its simplex constants per node,
// with reciprocals of the hardware (one ulp in FP32, two Newton steps in
// FP64) in place of correctly rounded divisions, and its vertex lengths with
// the same square root; all per-node-metric kernels (the metric engine, the
// per-node plane kernel, the batched solve) do so alike and give each other's
// bits.  Against correctly rounded constants a candidate moves by a few ulps,
// far inside the plane bound's shrink (plane_gap).
template <typename R>
__device__ __forceinline__ R lat_rcp(R x);
template <>
__device__ __forceinline__ float lat_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <>
__device__ __forceinline__ double lat_rcp<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
template <typename R>
__device__ __forceinline__ R lat_fma(R a, R b, R c);
template <>
__device__ __forceinline__ float lat_fma<float>(float a, float b, float c) { return fmaf(a, b, c); }
template <>
__device__ __forceinline__ double lat_fma<double>(double a, double b, double c) {
  return fma(a, b, c);
}
// the square root of a lattice length or a candidate's radicand
__device__ __forceinline__ double lat_sqrt(double x) { return sqrt_rn(x); }
__device__ __forceinline__ float lat_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <typename R, bool kGap>
__device__ __forceinline__ void ellip_edge_lat(R &best, R t0, R t1, R s2, R P, R l0, R kap,
                                               R gap) {
  if (!may_win<R, kGap>(rmin(t0, t1), gap, best) || !(reached(t0) && reached(t1))) return;
  const R d = t1 - t0;
  const R mu = P * d;
  const R r = lat_fma(-d, mu, s2);
  if (!(r > R(0))) return;
  const R root = lat_sqrt(kap * r);
  const R rm = -root * mu;
  if (!(lat_fma(l0, r, rm) >= R(0))) return;           // foot point past t0
  if (!(lat_fma(l0 - R(1), r, rm) <= R(0))) return;    // and before t1
  const R c = lat_fma(d, l0, t0) + root;
  if (c < best) best = c;
}

template <typename R, bool kGap>
__device__ __forceinline__ void ellip_tri_lat(R &best, R t0, R t1, R t2, R s2, R P11, R P12,
                                              R P22, R l1, R l2, R kap, R gap) {
  if (!may_win<R, kGap>(rmin(t0, rmin(t1, t2)), gap, best) ||
      !(reached(t0) && reached(t1) && reached(t2)))
    return;
  const R d1 = t1 - t0, d2 = t2 - t0;
  const R m1 = lat_fma(P11, d1, P12 * d2);
  const R m2 = lat_fma(P12, d1, P22 * d2);
  const R r = lat_fma(-d2, m2, lat_fma(-d1, m1, s2));
  if (!(r > R(0))) return;
  const R root = lat_sqrt(kap * r);
  const R rm1 = -root * m1, rm2 = -root * m2;
  if (!(lat_fma(l1, r, rm1) >= R(0))) return;
  if (!(lat_fma(l2, r, rm2) >= R(0))) return;
  if (!(lat_fma(l1 + l2 - R(1), r, rm1 + rm2) <= R(0))) return;
  const R c = lat_fma(d2, l2, lat_fma(d1, l1, t0)) + root;
  if (c < best) best = c;
}

// The same candidates with the constants of a table (EllipConsts), for a
// constant metric.  FP64 runs the fused form above, which the oracle restates.
// FP32 keeps one rounding per operation: with the fused form the per-iteration
// max|dT| of a converged FP32 solve creeps at 1.0e-6 (512^3 blob speed, some
// 8 ulps of T), just above that workload's 1e-6 stop, where the unfused form
// stays at 3-4e-7, so the stop would come two iterations later.
template <typename R, bool kGap = false>
__device__ __forceinline__ void ellip_edge(R &best, R t0, R t1, R s2, const R *E, R gap = R(0)) {
  if constexpr (sizeof(R) == 8) {
    ellip_edge_lat<R, kGap>(best, t0, t1, s2, E[0], E[1], E[2], gap);
  } else {
    if (!may_win<R, kGap>(rmin(t0, t1), gap, best) || !(reached(t0) && reached(t1))) return;
    const R P = E[0], l0 = E[1], kap = E[2];
    const R d = t1 - t0;
    const R mu = P * d;
    const R r = s2 - d * mu;
    if (!(r > R(0))) return;
    const R root = lat_sqrt(kap * r);
    if (!(l0 * r - root * mu >= R(0))) return;
    if (!((l0 - R(1)) * r - root * mu <= R(0))) return;
    const R c = t0 + d * l0 + root;
    if (c < best) best = c;
  }
}
template <typename R, bool kGap = false>
__device__ __forceinline__ void ellip_tri(R &best, R t0, R t1, R t2, R s2, const R *Tr,
                                          R gap = R(0)) {
  if constexpr (sizeof(R) == 8) {
    ellip_tri_lat<R, kGap>(best, t0, t1, t2, s2, Tr[0], Tr[1], Tr[2], Tr[3], Tr[4], Tr[5], gap);
  } else {
    if (!may_win<R, kGap>(rmin(t0, rmin(t1, t2)), gap, best) ||
        !(reached(t0) && reached(t1) && reached(t2)))
      return;
    const R d1 = t1 - t0, d2 = t2 - t0;
    const R m1 = Tr[0] * d1 + Tr[1] * d2;
    const R m2 = Tr[1] * d1 + Tr[2] * d2;
    const R r = s2 - (d1 * m1 + d2 * m2);
    if (!(r > R(0))) return;
    const R root = lat_sqrt(Tr[5] * r);
    if (!(Tr[3] * r - root * m1 >= R(0))) return;
    if (!(Tr[4] * r - root * m2 >= R(0))) return;
    if (!((Tr[3] + Tr[4] - R(1)) * r - root * (m1 + m2) <= R(0))) return;
    const R c = t0 + (d1 * Tr[3] + d2 * Tr[4]) + root;
    if (c < best) best = c;
  }
}
// A one-point candidate t + s |X - Y|_N with a table's length, in the same form.
template <typename R>
__device__ __forceinline__ R ellip_vertex(R t, R s, R len) {
  if constexpr (sizeof(R) == 8) return lat_fma(s, len, t);
  else return t + s * len;
}

// p[0] = centre, p[1..8] = ring.  Returns the new (non-increasing) value.
template <typename R>
__device__ __forceinline__ R ellip_node_update(R told, const R p[9], R s,
                                               const EllipConsts<R> &K) {
  R best = told;
  const R s2 = s * s;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    if (reached(p[k])) best = rmin(ellip_vertex(p[k], s, K.vert[k]), best);
#pragma unroll
  for (int k = 0; k < 8; ++k) ellip_edge(best, p[0], p[1 + k], s2, K.edge[k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) ellip_edge(best, p[1 + k], p[1 + (k + 1) % 8], s2, K.edge[8 + k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) ellip_tri(best, p[0], p[1 + k], p[1 + (k + 1) % 8], s2, K.tri[k]);
  return best;
}

// Local-frame constants for inverse metric N (xx, yy, zz, xy, xz, yz in global
// axes), frame axes u = (x^, y^, z^) with signs sg, spacings (dxx, dyy, dzz).
// Used on the host for a constant metric and on the device per node.
template <typename R>
__host__ __device__ inline void ellip_constants(const R N[6], const int u[3],
                                                const R sg[3], const R hl[3],
                                                EllipConsts<R> &K) {
  const R Ng[3][3] = {{N[0], N[3], N[4]}, {N[3], N[1], N[5]}, {N[4], N[5], N[2]}};
  R NL[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) NL[i][j] = sg[i] * sg[j] * Ng[u[i]][u[j]];
  const R dxx = hl[0], dyy = hl[1], dzz = hl[2];
  R pos[9][3];
  pos[0][0] = R(0);
  pos[0][1] = R(0);
  pos[0][2] = R(0);
  for (int k = 0; k < 8; ++k) {
    pos[k + 1][0] = kRing[k][0] * dxx;
    pos[k + 1][1] = kRing[k][1] * dyy;
    pos[k + 1][2] = R(0);
  }
  const R X[3] = {R(0), R(0), dzz};
  auto ndot = [&](const R *p, const R *q) {
    return p[0] * (NL[0][0] * q[0] + NL[0][1] * q[1] + NL[0][2] * q[2]) +
           p[1] * (NL[1][0] * q[0] + NL[1][1] * q[1] + NL[1][2] * q[2]) +
           p[2] * (NL[2][0] * q[0] + NL[2][1] * q[1] + NL[2][2] * q[2]);
  };
  for (int k = 0; k < 9; ++k) {
    R D[3] = {X[0] - pos[k][0], X[1] - pos[k][1], X[2] - pos[k][2]};
    K.vert[k] = sqrt(ndot(D, D));
  }
  for (int k = 0; k < 16; ++k) {
    const int v0 = k < 8 ? 0 : 1 + (k - 8);
    const int v1 = k < 8 ? 1 + k : 1 + ((k - 8 + 1) % 8);
    R E[3], D[3];
    for (int i = 0; i < 3; ++i) {
      E[i] = pos[v1][i] - pos[v0][i];
      D[i] = X[i] - pos[v0][i];
    }
    const R g = ndot(E, E), b = ndot(E, D), c = ndot(D, D);
    const R P = R(1) / g;
    const R l0 = P * b;
    K.edge[k][0] = P;
    K.edge[k][1] = l0;
    K.edge[k][2] = c - b * l0;
  }
  for (int k = 0; k < 8; ++k) {
    const int v1 = 1 + k, v2 = 1 + (k + 1) % 8;
    const R *E1 = pos[v1], *E2 = pos[v2];
    const R g11 = ndot(E1, E1), g12 = ndot(E1, E2), g22 = ndot(E2, E2);
    const R b1 = ndot(E1, X), b2 = ndot(E2, X), c = ndot(X, X);
    const R det = g11 * g22 - g12 * g12;
    const R P11 = g22 / det, P12 = -g12 / det, P22 = g11 / det;
    const R l1 = P11 * b1 + P12 * b2, l2 = P12 * b1 + P22 * b2;
    K.tri[k][0] = P11;
    K.tri[k][1] = P12;
    K.tri[k][2] = P22;
    K.tri[k][3] = l1;
    K.tri[k][4] = l2;
    K.tri[k][5] = c - (b1 * l1 + b2 * l2);
  }
}

// ---------------------------------------------------------------------------
// Per-node metric.  Every patch vector is a lattice vector w = (a dxx, b dyy,
// c dzz) with small integers a, b, c, so w^T N w and v^T N w are integer
// combinations of six products of the local inverse metric with the spacings.
// A unit's constants are then a handful of FMAs and one division, and they are
// computed only for units that survive the pruning.
template <typename R>
struct LatticeN {
  R A, B, C, E, Fx, Fy;  // n11 dxx^2, n22 dyy^2, n33 dzz^2, n12 dxx dyy, n13 dxx dzz, n23 dyy dzz
};

template <typename R>
__device__ __forceinline__ R lat_dot(const LatticeN<R> &L, int a1, int b1, int c1, int a2,
                                     int b2, int c2) {
  return R(a1 * a2) * L.A + R(b1 * b2) * L.B + R(c1 * c2) * L.C + R(a1 * b2 + a2 * b1) * L.E +
         R(a1 * c2 + a2 * c1) * L.Fx + R(b1 * c2 + b2 * c1) * L.Fy;
}

// lat_dot with compile-time lattice vectors: the same six products summed in
// the same order, with the terms whose integer coefficient is zero left out.
// Those terms are exact zeros (the lattice products are finite), and adding a
// zero leaves a nonzero partial sum unchanged, so the result is lat_dot's
// (only the sign of an all-zero sum could differ, and no candidate depends on
// it).  Under IEEE rules the compiler may not drop 0 * x itself: lat_dot
// spends eleven instructions on every dot product, this form one to three for
// the patch's lattice vectors.
template <typename R, int k>
__device__ __forceinline__ R lat_term(R x) {
  if constexpr (k == 1) return x;
  else if constexpr (k == -1) return -x;
  else return R(k) * x;
}
template <typename R, int k, int kBefore>
__device__ __forceinline__ void lat_add(R &acc, R x) {
  if constexpr (k != 0) {
    if constexpr (kBefore == 0) acc = lat_term<R, k>(x);
    else acc = acc + lat_term<R, k>(x);
  }
}
template <int a1, int b1, int c1, int a2, int b2, int c2, typename R>
__device__ __forceinline__ R lat_dot_c(const LatticeN<R> &L) {
  constexpr int kA = a1 * a2, kB = b1 * b2, kC = c1 * c2, kE = a1 * b2 + a2 * b1,
                kFx = a1 * c2 + a2 * c1, kFy = b1 * c2 + b2 * c1;
  constexpr int n1 = kA != 0, n2 = n1 + (kB != 0), n3 = n2 + (kC != 0), n4 = n3 + (kE != 0),
                n5 = n4 + (kFx != 0), n6 = n5 + (kFy != 0);
  static_assert(n6 > 0, "a lattice dot product with no term");
  R acc = R(0);
  lat_add<R, kA, 0>(acc, L.A);
  lat_add<R, kB, n1>(acc, L.B);
  lat_add<R, kC, n2>(acc, L.C);
  lat_add<R, kE, n3>(acc, L.E);
  lat_add<R, kFx, n4>(acc, L.Fx);
  lat_add<R, kFy, n5>(acc, L.Fy);
  return acc;
}

// The six products P_ab = (N_ab h_a) h_b of N with the spacings of the global
// axes a <= b, in the symmetric storage order (xx, yy, zz, xy, xz, yz).  A
// frame's lattice products are these up to sign, and every per-node-metric
// kernel forms them this way (the metric fields hold them), so all of them
// give each other's bits.
template <typename R>
__host__ __device__ inline void lattice_products(const R N[6], const R h[3], R P[6]) {
  P[0] = (N[0] * h[0]) * h[0];
  P[1] = (N[1] * h[1]) * h[1];
  P[2] = (N[2] * h[2]) * h[2];
  P[3] = (N[3] * h[0]) * h[1];
  P[4] = (N[4] * h[0]) * h[2];
  P[5] = (N[5] * h[1]) * h[2];
}

// Lattice products of a frame (axes u, signs sg, spacings hl in frame order)
// from N: sg_i sg_j P_(u_i u_j).
template <typename R>
__device__ __forceinline__ LatticeN<R> lattice_of(const R N[6], const int u[3], const R sg[3],
                                                  const R hl[3]) {
  const R Ng[3][3] = {{N[0], N[3], N[4]}, {N[3], N[1], N[5]}, {N[4], N[5], N[2]}};
  auto nl = [&](int i, int j) {
    const R n = Ng[u[i]][u[j]];
    const R p = u[i] <= u[j] ? (n * hl[i]) * hl[j] : (n * hl[j]) * hl[i];
    return (sg[i] * sg[j]) * p;
  };
  LatticeN<R> L;
  L.A = nl(0, 0);
  L.B = nl(1, 1);
  L.C = nl(2, 2);
  L.E = nl(0, 1);
  L.Fx = nl(0, 2);
  L.Fy = nl(1, 2);
  return L;
}

// Lower bound, for the pruning test, of s |X - Y|_N over the points Y of the
// upwind plane: the minimum of |w|_N^2 over w = (a dxx, b dyy, dzz), a and b
// real, is the Schur complement C - (B Fx^2 - 2 E Fx Fy + A Fy^2) / (A B - E^2).
// It is shrunk to leave room for the rounding of a candidate: near a grazing
// simplex the square root of the rounded s^2 - delta.mu moves a candidate by up
// to about 5 sqrt(eps) s h sqrt(lambda_max(N)): in FP32 2e-3 of the bound for
// a metric of condition number 1 and 0.1 at about 3000, in FP64 1e-3 of the
// bound at 2e8.  So FP32 keeps 90 % of the bound and FP64 99.9 %.  Ill-conditioned cases (the complement below
// 1e-3 of C, or a nearly singular in-plane block) get no bound.
template <typename R>
__device__ __forceinline__ R plane_gap(const LatticeN<R> &L, R s) {
  const R det = L.A * L.B - L.E * L.E;
  if (!(det > R(1e-6) * L.A * L.B)) return R(0);
  // (the hardware reciprocal and square root: a bound is all that is needed,
  // and their few ulps lie far inside the shrink)
  const R q =
      L.C - (L.B * L.Fx * L.Fx - R(2) * L.E * L.Fx * L.Fy + L.A * L.Fy * L.Fy) * lat_rcp(det);
  if (!(q > R(1e-3) * L.C)) return R(0);
  return (sizeof(R) == 8 ? R(0.999) : R(0.9)) * s * lat_sqrt(q);
}

// Unit U of the node (triangle U, spoke U, rim U) with constants from L.
template <typename R, int U, bool kGap = false>
__device__ __forceinline__ void ellip_unit_lattice(R &best, const R p[9], R s2,
                                                   const LatticeN<R> &L, R gap = R(0)) {
  constexpr int k1 = 1 + U, k2 = 1 + (U + 1) % 8;
  constexpr int a1 = kRing[U][0], b1 = kRing[U][1];
  constexpr int a2 = kRing[(U + 1) % 8][0], b2 = kRing[(U + 1) % 8][1];
  const R t0 = p[0], t1 = p[k1], t2 = p[k2];
  if (!may_win<R, kGap>(rmin(t0, rmin(t1, t2)), gap, best)) return;  // nothing here can win
  {  // spoke centre -> ring U: E = (a1, b1, 0), D = (0, 0, 1)
    const R g = lat_dot_c<a1, b1, 0, a1, b1, 0>(L);
    const R b = lat_dot_c<a1, b1, 0, 0, 0, 1>(L);
    const R P = lat_rcp(g), l0 = P * b;
    ellip_edge_lat<R, kGap>(best, t0, t1, s2, P, l0, lat_fma(-b, l0, L.C), gap);
  }
  {  // rim ring U -> ring U+1: E = (a2-a1, b2-b1, 0), D = (-a1, -b1, 1)
    constexpr int ea = a2 - a1, eb = b2 - b1;
    const R g = lat_dot_c<ea, eb, 0, ea, eb, 0>(L);
    const R b = lat_dot_c<ea, eb, 0, -a1, -b1, 1>(L);
    const R c = lat_dot_c<-a1, -b1, 1, -a1, -b1, 1>(L);
    const R P = lat_rcp(g), l0 = P * b;
    ellip_edge_lat<R, kGap>(best, t1, t2, s2, P, l0, lat_fma(-b, l0, c), gap);
  }
  {  // triangle centre, ring U, ring U+1
    const R g11 = lat_dot_c<a1, b1, 0, a1, b1, 0>(L), g22 = lat_dot_c<a2, b2, 0, a2, b2, 0>(L);
    const R g12 = lat_dot_c<a1, b1, 0, a2, b2, 0>(L);
    const R q1 = lat_dot_c<a1, b1, 0, 0, 0, 1>(L), q2 = lat_dot_c<a2, b2, 0, 0, 0, 1>(L);
    const R rdet = lat_rcp(lat_fma(g11, g22, -g12 * g12));
    const R P11 = g22 * rdet, P12 = -g12 * rdet, P22 = g11 * rdet;
    const R l1 = lat_fma(P11, q1, P12 * q2), l2 = lat_fma(P12, q1, P22 * q2);
    ellip_tri_lat<R, kGap>(best, t0, t1, t2, s2, P11, P12, P22, l1, l2,
                           lat_fma(-q2, l2, lat_fma(-q1, l1, L.C)), gap);
  }
}

// The nine one-point candidates of a node with a per-node metric.
template <typename R, int K>
__device__ __forceinline__ void ellip_vertex_lattice(R &best, const R p[9], R s,
                                                     const LatticeN<R> &L) {
  constexpr int a = K == 0 ? 0 : kRing[K - 1][0], b = K == 0 ? 0 : kRing[K - 1][1];
  if (reached(p[K]))
    best = rmin(lat_fma(s, lat_sqrt(lat_dot_c<-a, -b, 1, -a, -b, 1>(L)), p[K]), best);
  if constexpr (K < 8) ellip_vertex_lattice<R, K + 1>(best, p, s, L);
}
template <typename R>
__device__ __forceinline__ void ellip_vertices_lattice(R &best, const R p[9], R s,
                                                       const LatticeN<R> &L) {
  ellip_vertex_lattice<R, 0>(best, p, s, L);
}

// Whole node with a per-node metric (all eight units; the nine one-point
// candidates first so the pruning starts low).
template <typename R>
__device__ __forceinline__ R ellip_node_lattice(R told, const R p[9], R s, const LatticeN<R> &L) {
  R best = told;
  ellip_vertices_lattice<R>(best, p, s, L);
  const R s2 = s * s;
  ellip_unit_lattice<R, 0>(best, p, s2, L);
  ellip_unit_lattice<R, 1>(best, p, s2, L);
  ellip_unit_lattice<R, 2>(best, p, s2, L);
  ellip_unit_lattice<R, 3>(best, p, s2, L);
  ellip_unit_lattice<R, 4>(best, p, s2, L);
  ellip_unit_lattice<R, 5>(best, p, s2, L);
  ellip_unit_lattice<R, 6>(best, p, s2, L);
  ellip_unit_lattice<R, 7>(best, p, s2, L);
  return best;
}

// M = R diag(1, ra, rb)^2 R^T with R = Rz(phi) Ry(theta); returns N = M^-1 =
// R diag(1, 1/ra^2, 1/rb^2) R^T directly (no 3x3 inverse needed).
template <typename R>
__host__ __device__ inline void tilted_inverse_metric(R theta, R phi, R ra, R rb, R N[6]) {
  const R ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
  const R Q[3][3] = {{cp * ct, -sp, cp * st}, {sp * ct, cp, sp * st}, {-st, R(0), ct}};
  const R w[3] = {R(1), R(1) / (ra * ra), R(1) / (rb * rb)};
  R F[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R acc = R(0);
      for (int k = 0; k < 3; ++k) acc += Q[i][k] * w[k] * Q[j][k];
      F[i][j] = acc;
    }
  N[0] = F[0][0];
  N[1] = F[1][1];
  N[2] = F[2][2];
  N[3] = F[0][1];
  N[4] = F[0][2];
  N[5] = F[1][2];
}

}  // namespace fold3d
