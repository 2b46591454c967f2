// fold_fast.h — the fold-model node update of fold_local.h with its
// per-direction invariants hoisted out of the node loop.
//
// In the reference's local solvers (AN_ISO_1_OpenACC.c:59-175) every
// parameter except the ten patch values is fixed per sweep direction and per
// sign of the target node (a flips where T < 0):
//   * each edge call site has fixed (ax, ay, az, dxx, dzz), so its one-point
//     candidate xt + len2/c (:99-113) is xt + a constant;
//   * the one-point update from the centre (:214-228) is st + a constant;
//   * the rotated-frame parameters of the diagonal edges and of the triangle
//     sub-edges ((ax*dxx + ay*dyy)/dxy, (ax*dzz + az*dxx)/dxz, ...) are constants;
//   * every division by one of those constants (x/dxx, x/(ga - az^2), x/ga) is
//     computed as a reciprocal multiply with two fused residual corrections,
//     which returns the correctly rounded quotient (Markstein's theorem: once
//     q is within one ulp of x/d, q + (x - q d) * RN(1/d) rounds to RN(x/d)),
//     i.e. the same bits as the IEEE division.
// Everything else keeps the reference's operation order, so results stay
// bit-identical to fold_local.h: to the reference in FP64, and to the literal
// kernel run in FP32 (the constants are then computed in float, with the
// same expressions, and the square root is the correctly rounded one).
//
// One more early exit, beyond the reference's: every candidate of a simplex
// in the upwind plane is its foot point's interpolated value (>= the
// simplex's smallest vertex) plus the travel time tau(X - Y) from the foot
// point Y to the target X.  For F|p| + a.p = 1, tau(w) = max p.w over
// H(p) <= 1, so by minimax the least travel time to X from an affine set S
// (a line or a plane through Y0) is the largest p.(X - Y0) over the p with
// H(p) <= 1 orthogonal to S: from the plane z = 0 to
// (0, 0, dzz) it is dzz / (F + az) (p on the z axis); from an edge's line (an
// edge site's own frame: the line is the x axis) it is dzz / (sqrt(F^2 - ay^2)
// + az) (p in the y-z plane).  A triangle (with its sub-edge) or an edge
// whose smallest vertex plus its bound (shrunk to 99.9 % in FP64, 90 % in
// FP32, against the rounding of a candidate near a grazing characteristic) is
// not below the running value by 16 ulps cannot lower it, and is skipped.  The
// literal kernel (fold_local.h) does not skip; the tests hold both to the
// reference's bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "fold_local.h"

namespace fold3d {

// x / d with r = RN(1/d): correctly rounded for finite x, d and normal results.
__device__ __forceinline__ double div_const(double x, double d, double r) {
  double q = x * r;
  double e = fma(-q, d, x);
  q = fma(e, r, q);
  e = fma(-q, d, x);
  return fma(e, r, q);
}
__device__ __forceinline__ float div_const(float x, float d, float r) {
  float q = x * r;
  float e = fmaf(-q, d, x);
  q = fmaf(e, r, q);
  e = fmaf(-q, d, x);
  return fmaf(e, r, q);
}

template <typename R>
__device__ __forceinline__ R fast_sqrt(R x);
template <>
__device__ __forceinline__ double fast_sqrt<double>(double x) { return sqrt_rn(x); }
// FP32: the correctly rounded square root, as sqrt() of the literal kernel
// (fold_local.h) compiles to, so both give the same bits.
template <>
__device__ __forceinline__ float fast_sqrt<float>(float x) { return __builtin_sqrtf(x); }

// The smaller of two magnitudes of the field or two candidates.  These are
// non-negative, never -0 (magnitudes are fabs() of the field; an accepted
// candidate exceeds the smallest vertex) and never NaN for a field without
// NaN, so fmin gives the value rmin gives, in one v_min instruction where
// rmin's compare and two selects take three.  (In the foot-point test
// min(fx, fy) a NaN operand makes the test fail either way.)
__device__ __forceinline__ double vmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }

// One edge call site (FoldEdge2Points with fixed model/spacing arguments).
template <typename R>
struct EdgeSite {
  R ax, az, dxx, dzz;
  R ga, rga;        // F^2 - ay^2 and its reciprocal
  R gaz, rgaz;      // ga - az^2 and its reciprocal
  R rdxx;           // 1/dxx
  R k1;             // one-point constant len2/c of the far end
  R lgap;           // shrunk least travel time from the edge's line to the target
  int ax_pos;       // ax > 0
  int k1_ok;        // the one-point candidate exists
};

// One triangle call site (FoldSurf3Points with fixed arguments) and its sub-edge.
template <typename R>
struct TriSite {
  R ax, ay, dxx, dyy, rdxx, rdyy;
  int ax_pos, ay_pos;
  int sub_pos;      // ax*dzz + az*dxx > 0
};

// Everything one sign of one direction needs.
template <typename R>
struct FoldSign {
  R F, az, dzz;
  R gap;            // shrunk dzz / (F + az): least travel time from the upwind plane
  R ga, rga;        // triangles: F^2 - az^2 (shared by all eight) and reciprocal
  R k0;             // centre one-point constant dzz^2/d
  int k0_ok;
  TriSite<R> tri[8];
  EdgeSite<R> edge[16];  // 0..7 the node's edges, 8..15 the triangles' sub-edges
};

template <typename R>
struct FoldFast {
  FoldSign<R> sgn[2];  // [0]: T_target >= 0 (a as given), [1]: T_target < 0 (a negated)
};

// Both candidate functions leave at once when tnew is at most the smallest of
// their vertices (the reference tests tnew < smallest): a candidate is taken
// only above its smallest vertex, and the one-point values lie above their
// vertex, so none can go below tnew.  On ties this skips work the reference
// does and rejects; in particular it skips the unreached regions, where the
// target and its whole patch hold MAXT.
// Constants held in scalar registers from one point on.  The node update
// reads its per-direction constants with wave-uniform addresses (scalar
// loads); left alone, the compiler sinks each load into the branch that first
// uses it, and a wave then waits on one scalar-cache round trip per branch.
// Naming a group of constants in one empty asm statement issues their loads
// together and waits once.
template <typename R>
__device__ __forceinline__ void pin_edge_body(const EdgeSite<R> &e) {
  asm volatile("" ::"s"(e.dxx), "s"(e.rdxx), "s"(e.ax), "s"(e.az), "s"(e.ga), "s"(e.gaz),
               "s"(e.rgaz), "s"(e.dzz), "s"(e.rga), "s"(e.k1));
}
template <typename R>
__device__ __forceinline__ void pin_tri_body(const FoldSign<R> &S, const TriSite<R> &t,
                                             const EdgeSite<R> &sub) {
  asm volatile("" ::"s"(t.dxx), "s"(t.rdxx), "s"(t.dyy), "s"(t.rdyy), "s"(t.ax), "s"(t.ay),
               "s"(S.ga), "s"(S.rga), "s"(sub.ax_pos), "s"(sub.lgap));
}
template <typename R>
__device__ __forceinline__ void pin_unit_gate(const FoldSign<R> &S, const TriSite<R> &t,
                                              const EdgeSite<R> &e) {
  asm volatile("" ::"s"(t.ax_pos), "s"(t.ay_pos), "s"(t.sub_pos), "s"(S.gap), "s"(e.ax_pos),
               "s"(e.lgap), "s"(e.k1_ok));
}

template <typename R>
__device__ __forceinline__ void fast_edge(R &tnew, R st, R xt, R F, const EdgeSite<R> &e) {
  if (tnew <= vmin(xt, st) || (e.ax_pos && xt > st)) return;
  if (!(vmin(xt, st) + e.lgap < tnew * prune_slack<R>())) return;  // cannot go below tnew
  pin_edge_body(e);
  const R gx = div_const(xt - st, e.dxx, e.rdxx);
  const R q = R(1) - gx * e.ax;
  const R disc = e.ga * (q * q - gx * gx * e.gaz);
  if (R(0) < disc) {
    const R gz = div_const(-e.az * q + fast_sqrt(disc), e.gaz, e.rgaz);
    const R cand = st + e.dzz * gz;
    if (cand < tnew && cand > vmin(st, xt)) {
      R t = (gx * gx + gz * gz) * F;
      t = t * F;
      const R g = fast_sqrt(div_const(t, e.ga, e.rga));
      const R foot = -e.dzz * (gx + e.ax * g) / (gz + e.az * g);
      if (R(0) < foot && foot < e.dxx) tnew = cand;
    }
  }
  if (tnew > xt && e.k1_ok) tnew = vmin(tnew, xt + e.k1);
}

template <typename R>
__device__ __forceinline__ void fast_tri(R &tnew, R st, R xt, R yt, const FoldSign<R> &S,
                                         const TriSite<R> &t, const EdgeSite<R> &sub) {
  if (tnew <= vmin(st, vmin(xt, yt)) || (t.ax_pos && xt > st) || (t.ay_pos && yt > xt)) return;
  if (!(vmin(st, vmin(xt, yt)) + S.gap < tnew * prune_slack<R>())) return;  // nor its sub-edge
  pin_tri_body(S, t, sub);
  const R F = S.F, az = S.az, dzz = S.dzz;
  const R gx = div_const(xt - st, t.dxx, t.rdxx);
  const R gy = div_const(yt - xt, t.dyy, t.rdyy);
  const R g2 = gx * gx + gy * gy;
  const R al = R(1) - t.ax * gx - t.ay * gy;
  const R disc = al * al - g2 * S.ga;
  if (disc > R(0)) {
    const R gz = div_const(-az * al + F * fast_sqrt(disc), S.ga, S.rga);
    const R cand = st + dzz * gz;
    if (cand < tnew && cand > vmin(vmin(st, xt), yt)) {
      const R g = fast_sqrt(g2 + gz * gz);
      const R s = -dzz / (F * gz + az * g);
      const R fx = s * (F * gx + t.ax * g);
      const R fy = s * (F * gy + t.ay * g);
      if (R(0) <= vmin(fx, fy) && fy * t.dxx <= fx * t.dyy && fx <= t.dxx) tnew = cand;
    }
  }
  if (tnew > vmin(xt, yt) && (!(t.sub_pos && xt > yt))) fast_edge(tnew, xt, yt, F, sub);
}

// Same result as fold_node_update(), bit for bit.
template <typename R>
__device__ __forceinline__ R fold_node_fast(R told, const Patch<R> &p, const FoldSign<R> &S) {
  R tnew = fabs(told);
  const R st = p.c;
  if (st < tnew && S.k0_ok) tnew = vmin(tnew, st + S.k0);
  if (!(tnew < st && tnew < vmin(p.px, p.py) && tnew < vmin(p.pxpy, p.pxmy) &&
        tnew < vmin(p.mxmy, p.mxpy))) {
    // (centre, first, second) vertices of the eight triangles, reference order
    fast_tri(tnew, st, p.px, p.pxpy, S, S.tri[0], S.edge[8]);
    fast_tri(tnew, st, p.px, p.pxmy, S, S.tri[1], S.edge[9]);
    fast_tri(tnew, st, p.mx, p.mxmy, S, S.tri[2], S.edge[10]);
    fast_tri(tnew, st, p.mx, p.mxpy, S, S.tri[3], S.edge[11]);
    fast_tri(tnew, st, p.my, p.mxmy, S, S.tri[4], S.edge[12]);
    fast_tri(tnew, st, p.my, p.pxmy, S, S.tri[5], S.edge[13]);
    fast_tri(tnew, st, p.py, p.pxpy, S, S.tri[6], S.edge[14]);
    fast_tri(tnew, st, p.py, p.mxpy, S, S.tri[7], S.edge[15]);
    fast_edge(tnew, st, p.px, S.F, S.edge[0]);
    fast_edge(tnew, st, p.pxpy, S.F, S.edge[1]);
    fast_edge(tnew, st, p.py, S.F, S.edge[2]);
    fast_edge(tnew, st, p.mxpy, S.F, S.edge[3]);
    fast_edge(tnew, st, p.mx, S.F, S.edge[4]);
    fast_edge(tnew, st, p.mxmy, S.F, S.edge[5]);
    fast_edge(tnew, st, p.my, S.F, S.edge[6]);
    fast_edge(tnew, st, p.pxmy, S.F, S.edge[7]);
  }
  return ((told > R(0)) ? R(1) : R(-1)) * tnew;
}

// The node update cut into eight units, each one triangle (with its sub-edge)
// and one of the node's edges.  The minimum over the units' candidates and
// the centre one-point value equals fold_node_fast() (the candidates may be
// taken in any order), so a block can let different waves evaluate different
// units of the same nodes and take the minimum at the end.
template <typename R>
__device__ __forceinline__ R fold_start(R told, R st, const FoldSign<R> &S) {
  R tnew = fabs(told);
  if (st < tnew && S.k0_ok) tnew = vmin(tnew, st + S.k0);
  return tnew;
}

// The reference's gate (AN_ISO_1_OpenACC.c:231): with it closed the node
// keeps its one-point value.
template <typename R>
__device__ __forceinline__ bool fold_gate_open(R tnew, const Patch<R> &p) {
  return !(tnew < p.c && tnew < vmin(p.px, p.py) && tnew < vmin(p.pxpy, p.pxmy) &&
           tnew < vmin(p.mxmy, p.mxpy));
}

// Unit u (uniform across the wave): triangle u with sub-edge 8+u, and edge E(u)
// with E = {px, pxmy, mx, mxpy, mxmy, my, pxpy, py}.
template <typename R>
__device__ __forceinline__ void fold_unit(R &tnew, const Patch<R> &p, const FoldSign<R> &S,
                                          int u) {
  const R st = p.c;
  {  // the unit's gate constants, loaded together (u is a compile-time constant)
    constexpr int kEdgeOf[8] = {0, 7, 4, 3, 5, 6, 1, 2};
    pin_unit_gate(S, S.tri[u], S.edge[kEdgeOf[u]]);
  }
  switch (u) {
    case 0:
      fast_tri(tnew, st, p.px, p.pxpy, S, S.tri[0], S.edge[8]);
      fast_edge(tnew, st, p.px, S.F, S.edge[0]);
      break;
    case 1:
      fast_tri(tnew, st, p.px, p.pxmy, S, S.tri[1], S.edge[9]);
      fast_edge(tnew, st, p.pxmy, S.F, S.edge[7]);
      break;
    case 2:
      fast_tri(tnew, st, p.mx, p.mxmy, S, S.tri[2], S.edge[10]);
      fast_edge(tnew, st, p.mx, S.F, S.edge[4]);
      break;
    case 3:
      fast_tri(tnew, st, p.mx, p.mxpy, S, S.tri[3], S.edge[11]);
      fast_edge(tnew, st, p.mxpy, S.F, S.edge[3]);
      break;
    case 4:
      fast_tri(tnew, st, p.my, p.mxmy, S, S.tri[4], S.edge[12]);
      fast_edge(tnew, st, p.mxmy, S.F, S.edge[5]);
      break;
    case 5:
      fast_tri(tnew, st, p.my, p.pxmy, S, S.tri[5], S.edge[13]);
      fast_edge(tnew, st, p.my, S.F, S.edge[6]);
      break;
    case 6:
      fast_tri(tnew, st, p.py, p.pxpy, S, S.tri[6], S.edge[14]);
      fast_edge(tnew, st, p.pxpy, S.F, S.edge[1]);
      break;
    default:
      fast_tri(tnew, st, p.py, p.mxpy, S, S.tri[7], S.edge[15]);
      fast_edge(tnew, st, p.py, S.F, S.edge[2]);
      break;
  }
}

// ---------------------------------------------------------------------------
// Construction of the constants, in the arithmetic of R (IEEE, no
// contraction, correctly rounded division and square root on the host and on
// the device alike) with the expressions of fold_local.h, so the hoisted
// values carry the bits the node loop of the literal kernel computes in the
// same precision.  Only the pruning bounds (gap, lgap) are computed in
// double: they are conservative bounds, not results.  Host and device code
// share these functions (the allocation-free solve builds its table on the
// device).

// Least travel time from the x axis to (0, 0, dzz) in an edge site's frame:
// dzz max_phi sin(phi) / (F + |ay| cos(phi) + az sin(phi)); the derivative
// vanishes only at cos(phi) = -|ay| / F, which gives dzz / (sqrt(F^2 - ay^2) +
// az) (>= the plane bound dzz / (F + az); positive because F > |a|).
__host__ __device__ inline double edge_line_bound(double F, double ay, double az, double dzz) {
  return dzz / (std::sqrt(F * F - ay * ay) + az);
}

// Local frame of direction `dir` from the global model (F, a) and spacings
// h, in R as the literal kernel receives it; DXYP/DXZP/DYZP of the reference
// driver (AN_ISO_1_OpenACC.c:1159-1161), lower axis first.
__host__ __device__ inline double face_diagonal(const double h[3], int u, int v) {
  const int lo = u < v ? u : v, hi = u < v ? v : u;
  return std::sqrt(h[lo] * h[lo] + h[hi] * h[hi]);
}
template <typename R>
__host__ __device__ inline FoldFrame<R> fold_frame(int dir, double F, const double a[3],
                                                   const double h[3]) {
  const Frame fr = kFrames[dir];
  FoldFrame<R> f;
  f.F = (R)F;
  f.ax = (R)(fr.sx * a[fr.ux]);
  f.ay = (R)(fr.sy * a[fr.uy]);
  f.az = (R)(fr.s * a[fr.m]);
  f.dzz = (R)h[fr.m];
  f.dxx = (R)h[fr.ux];
  f.dyy = (R)h[fr.uy];
  f.dxy = (R)face_diagonal(h, fr.ux, fr.uy);
  f.dxz = (R)face_diagonal(h, fr.ux, fr.m);
  f.dyz = (R)face_diagonal(h, fr.uy, fr.m);
  return f;
}

// Pruning bounds are shrunk by this factor: 99.9% in FP64, 90% in FP32.
template <typename R>
__host__ __device__ inline double prune_shrink() { return sizeof(R) == 8 ? 0.999 : 0.9; }

template <typename R>
__host__ __device__ inline void edge_site_consts(R F, R ax, R ay, R az, R dxx, R dzz,
                                                 EdgeSite<R> &e) {
  e.ax = ax;
  e.az = az;
  e.dxx = dxx;
  e.dzz = dzz;
  e.ga = F * F - ay * ay;
  e.rga = R(1) / e.ga;
  e.gaz = e.ga - az * az;
  e.rgaz = R(1) / e.gaz;
  e.rdxx = R(1) / dxx;
  e.lgap = (R)(prune_shrink<R>() * edge_line_bound(F, ay, az, dzz));
  e.ax_pos = ax > R(0);
  R c = (-dxx * ax + dzz * az);
  const R len2 = dxx * dxx + dzz * dzz;
  const R d = (F * F - ax * ax - ay * ay - az * az) * len2 + c * c;
  e.k1_ok = 0;
  e.k1 = R(0);
  if (d > R(0)) {
    c = c + std::sqrt(d);
    if (c > R(0)) {
      e.k1_ok = 1;
      e.k1 = len2 / c;
    }
  }
}

// FoldFrame f holds (F, ax, ay, az, dzz, dxx, dyy, dxy, dxz, dyz) of the
// direction in R, as the literal kernel receives it.
template <typename R>
__host__ __device__ inline void fold_sign_consts(const FoldFrame<R> &f, R sgn, FoldSign<R> &S) {
  const R F = f.F;
  // a flips on the negative side exactly as `a *= -1` does
  const R ax = f.ax * sgn, ay = f.ay * sgn, az = f.az * sgn;
  const R dzz = f.dzz, dxx = f.dxx, dyy = f.dyy, dxy = f.dxy, dxz = f.dxz, dyz = f.dyz;
  S.F = F;
  S.az = az;
  S.dzz = dzz;
  S.gap = (R)(prune_shrink<R>() * (double)dzz / ((double)F + (double)az));  // F > |a|
  S.ga = F * F - az * az;
  S.rga = R(1) / S.ga;
  R d = (F * F - (ax * ax + ay * ay)) * dzz * dzz;
  S.k0_ok = 0;
  S.k0 = R(0);
  if (d > R(0)) {
    d = az * dzz + std::sqrt(d);
    if (d > R(0)) {
      S.k0_ok = 1;
      S.k0 = dzz * dzz / d;
    }
  }
  // triangles (AN_ISO_1_OpenACC.c:233-240): (ax_t, ay_t, dxx_t, dyy_t, dxz_t)
  const R tp[8][5] = {
      {ax, ay, dxx, dyy, dxz},   {ax, -ay, dxx, dyy, dxz},  {-ax, -ay, dxx, dyy, dxz},
      {-ax, ay, dxx, dyy, dxz},  {-ay, -ax, dyy, dxx, dyz}, {-ay, ax, dyy, dxx, dyz},
      {ay, ax, dyy, dxx, dyz},   {ay, -ax, dyy, dxx, dyz}};
  for (int k = 0; k < 8; ++k) {
    TriSite<R> &t = S.tri[k];
    t.ax = tp[k][0];
    t.ay = tp[k][1];
    t.dxx = tp[k][2];
    t.dyy = tp[k][3];
    t.rdxx = R(1) / t.dxx;
    t.rdyy = R(1) / t.dyy;
    t.ax_pos = t.ax > R(0);
    t.ay_pos = t.ay > R(0);
    t.sub_pos = t.ax * dzz + az * t.dxx > R(0);
    const R txz = tp[k][4];
    // sub-edge xt -> yt in the rotated frame (:170-174)
    edge_site_consts<R>(F, t.ay, (t.ax * dzz + az * t.dxx) / txz, (-t.ax * t.dxx + az * dzz) / txz,
                      t.dyy, txz, S.edge[8 + k]);
  }
  // the node's own edges (:242-253), order px, pxpy, py, mxpy, mx, mxmy, my, pxmy
  edge_site_consts<R>(F, ax, ay, az, dxx, dzz, S.edge[0]);
  edge_site_consts<R>(F, (ax * dxx + ay * dyy) / dxy, (-ax * dyy + ay * dxx) / dxy, az, dxy, dzz,
                    S.edge[1]);
  edge_site_consts<R>(F, ay, ax, az, dyy, dzz, S.edge[2]);
  edge_site_consts<R>(F, -(ax * dxx - ay * dyy) / dxy, (ax * dyy + ay * dxx) / dxy, az, dxy, dzz,
                    S.edge[3]);
  edge_site_consts<R>(F, -ax, ay, az, dxx, dzz, S.edge[4]);
  edge_site_consts<R>(F, -(ax * dxx + ay * dyy) / dxy, (-ax * dyy + ay * dxx) / dxy, az, dxy, dzz,
                    S.edge[5]);
  edge_site_consts<R>(F, -ay, ax, az, dyy, dzz, S.edge[6]);
  edge_site_consts<R>(F, (ax * dxx - ay * dyy) / dxy, (ax * dyy + ay * dxx) / dxy, az, dxy, dzz,
                    S.edge[7]);
}

// Both signs of direction `dir`.
template <typename R>
__host__ __device__ inline void fold_fast_consts(int dir, double F, const double a[3],
                                                 const double h[3], FoldFast<R> &out) {
  const FoldFrame<R> f = fold_frame<R>(dir, F, a, h);
  fold_sign_consts<R>(f, R(1), out.sgn[0]);
  fold_sign_consts<R>(f, R(-1), out.sgn[1]);
}

}  // namespace fold3d
