// fold_local.h — device local solvers of the fold model
//     F*|grad T| + a.grad T = 1
// for one target node above a 3x3 upwind patch (reference
// Stencil-OpenACC/AN_ISO_1_OpenACC.c:59-175 for the edge/triangle solvers and
// :195-326 for the per-node update).
//
// The FP64 instantiation keeps the reference's operation order term by term
// and is built without FP contraction, so its results are bit-identical to a
// host build of the reference.  The FP32 instantiation runs the same formulas
// in single precision.
//
// The final value of a node is the minimum over the initial value, the
// one-point candidate and every *valid* edge/triangle candidate; validity
// never depends on the running minimum, and every early exit below only
// skips candidates that cannot be smaller than the running minimum.  So the
// candidates may be evaluated in any order without changing a bit.
#pragma once

#include <hip/hip_runtime.h>

namespace fold3d {

// Correctly rounded FP64 square root.  For 2^-767 <= x <= 2^1000 the
// refinement below of the hardware reciprocal square root (one coupled
// Newton step on (sqrt(x), 1/(2 sqrt(x))), then two residual corrections
// g + (x - g*g) * h) is the IEEE result, and it is the same operation sequence
// the compiler emits for sqrt() in that range; outside it (zero, tiny, huge,
// inf, NaN, negative) the compiler's full sqrt(), with its scaling and special
// cases, is used.  The test is wave-uniform, so a wave with every lane in range
// skips the scaling and special-case selects.
__device__ __forceinline__ double sqrt_refined(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

__device__ __forceinline__ double sqrt_rn(double x) {
  const bool in_range = x >= 0x1p-767 && x <= 0x1p+1000;
  if (__builtin_amdgcn_ballot_w64(!in_range) == 0) return sqrt_refined(x);
  return sqrt(x);
}

// 1 + 16 ulps: the pruning tests that add a travel-time bound to a simplex's
// smallest vertex allow this much for a candidate's own rounding.
template <typename R>
__device__ __forceinline__ R prune_slack();
template <>
__device__ __forceinline__ double prune_slack<double>() { return 1.0 + 16.0 * 2.220446049250313e-16; }
template <>
__device__ __forceinline__ float prune_slack<float>() { return 1.0f + 16.0f * 1.1920929e-7f; }

// Reference macro semantics: min(a,b) = a<b ? a : b (second operand on ties
// and NaN, AN_ISO_1_OpenACC.c:32-34); sign(v) = v>0 ? 1 : -1 (Fold3dPMM.h:32-34).
template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return (a < b) ? a : b; }

// Model and spacings of one sweep direction in its local frame: x^ and y^ span
// the upwind plane, z^ points from the upwind plane to the target plane.
template <typename R>
struct FoldFrame {
  R F, ax, ay, az;
  R dzz, dxx, dyy, dxy, dxz, dyz;
};

// Candidate from the edge st=(0,0) -> xt=(dxx,0) and the one-point candidate
// from xt, for the target at (0,0,dzz) (AN_ISO_1_OpenACC.c:59-114).
template <typename R>
__device__ __forceinline__ void edge_candidates(R &tnew, R st, R xt, R F, R ax,
                                                R ay, R az, R dxx, R dzz) {
  if (tnew < rmin(xt, st) || (ax > R(0) && xt > st)) return;
  R ga = F * F - ay * ay;
  R gx = (xt - st) / dxx;                       // in-plane gradient along the edge
  R q = R(1) - gx * ax;
  R disc = ga * (q * q - gx * gx * (ga - az * az));
  if (R(0) < disc) {
    R gz = (-az * q + sqrt(disc)) / (ga - az * az);  // out-of-plane gradient
    R cand = st + dzz * gz;
    if (cand < tnew && cand > rmin(st, xt)) {
      R g = sqrt((gx * gx + gz * gz) * F * F / ga);
      R foot = -dzz * (gx + ax * g) / (gz + az * g);  // where the characteristic meets the edge
      if (R(0) < foot && foot < dxx) tnew = cand;
    }
  }
  if (tnew > xt) {
    R b = (-dxx * ax + dzz * az);
    R len2 = dxx * dxx + dzz * dzz;
    R disc1 = (F * F - ax * ax - ay * ay - az * az) * len2 + b * b;
    if (disc1 > R(0)) {
      b = b + sqrt(disc1);
      if (b > R(0)) tnew = rmin(tnew, xt + len2 / b);
    }
  }
}

// Candidate from the triangle st=(0,0), xt=(dxx,0), yt=(dxx,dyy), then the
// edge xt -> yt solved in the frame rotated about y^ so that its z^ points
// from xt to the target (AN_ISO_1_OpenACC.c:119-175).
template <typename R>
__device__ __forceinline__ void triangle_candidates(R &tnew, R st, R xt, R yt,
                                                    R F, R ax, R ay, R az,
                                                    R dxx, R dyy, R dzz,
                                                    R dxz) {
  if (tnew < rmin(st, rmin(xt, yt)) || (ax > R(0) && xt > st) ||
      (ay > R(0) && yt > xt))
    return;
  R gx = (xt - st) / dxx;
  R gy = (yt - xt) / dyy;
  R g2 = gx * gx + gy * gy;
  R ga = F * F - az * az;
  R al = R(1) - ax * gx - ay * gy;
  R disc = al * al - g2 * ga;
  if (disc > R(0)) {
    R gz = (-az * al + F * sqrt(disc)) / ga;
    R cand = st + dzz * gz;
    if (cand < tnew && cand > rmin(rmin(st, xt), yt)) {
      R g = sqrt(g2 + gz * gz);
      R s = -dzz / (F * gz + az * g);
      R fx = s * (F * gx + ax * g);
      R fy = s * (F * gy + ay * g);
      if (R(0) <= rmin(fx, fy) && fy * dxx <= fx * dyy && fx <= dxx) tnew = cand;
    }
  }
  if (tnew > rmin(xt, yt) && (!(ax * dzz + az * dxx > R(0) && xt > yt)))
    edge_candidates(tnew, xt, yt, F, ay, (ax * dzz + az * dxx) / dxz,
                    (-ax * dxx + az * dzz) / dxz, dyy, dxz);
}

// The 3x3 upwind patch of one node, magnitudes, named in the local frame.
template <typename R>
struct Patch {
  R c;              // node below the target
  R px, py, mx, my; // +x^, +y^, -x^, -y^
  R pxpy, pxmy, mxpy, mxmy;
};

// New signed value of the target node (AN_ISO_1_OpenACC.c:195-326).
template <typename R>
__device__ __forceinline__ R fold_node_update(R told, const Patch<R> &p,
                                              const FoldFrame<R> &f) {
  const R F = f.F;
  const R dzz = f.dzz, dxx = f.dxx, dyy = f.dyy;
  const R dxy = f.dxy, dxz = f.dxz, dyz = f.dyz;
  R tnew = fabs(told);
  const R st = p.c;
  R ax = f.ax, ay = f.ay, az = f.az;
  if (told < R(0)) {  // the other side of the surface sees -a
    ax = -ax;
    ay = -ay;
    az = -az;
  }
  if (st < tnew) {
    R d = (F * F - (ax * ax + ay * ay)) * dzz * dzz;
    if (d > R(0)) {
      d = az * dzz + sqrt(d);
      if (d > R(0)) tnew = rmin(tnew, st + dzz * dzz / d);
    }
  }
  // Reference gate (:231): -x^ and -y^ neighbours take no part in it.
  if (!(tnew < st && tnew < rmin(p.px, p.py) && tnew < rmin(p.pxpy, p.pxmy) &&
        tnew < rmin(p.mxmy, p.mxpy))) {
    triangle_candidates(tnew, st, p.px, p.pxpy, F, ax, ay, az, dxx, dyy, dzz, dxz);
    triangle_candidates(tnew, st, p.px, p.pxmy, F, ax, -ay, az, dxx, dyy, dzz, dxz);
    triangle_candidates(tnew, st, p.mx, p.mxmy, F, -ax, -ay, az, dxx, dyy, dzz, dxz);
    triangle_candidates(tnew, st, p.mx, p.mxpy, F, -ax, ay, az, dxx, dyy, dzz, dxz);
    triangle_candidates(tnew, st, p.my, p.mxmy, F, -ay, -ax, az, dyy, dxx, dzz, dyz);
    triangle_candidates(tnew, st, p.my, p.pxmy, F, -ay, ax, az, dyy, dxx, dzz, dyz);
    triangle_candidates(tnew, st, p.py, p.pxpy, F, ay, ax, az, dyy, dxx, dzz, dyz);
    triangle_candidates(tnew, st, p.py, p.mxpy, F, ay, -ax, az, dyy, dxx, dzz, dyz);

    edge_candidates(tnew, st, p.px, F, ax, ay, az, dxx, dzz);
    edge_candidates(tnew, st, p.pxpy, F, (ax * dxx + ay * dyy) / dxy,
                    (-ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
    edge_candidates(tnew, st, p.py, F, ay, ax, az, dyy, dzz);
    edge_candidates(tnew, st, p.mxpy, F, -(ax * dxx - ay * dyy) / dxy,
                    (ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
    edge_candidates(tnew, st, p.mx, F, -ax, ay, az, dxx, dzz);
    edge_candidates(tnew, st, p.mxmy, F, -(ax * dxx + ay * dyy) / dxy,
                    (-ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
    edge_candidates(tnew, st, p.my, F, -ay, ax, az, dyy, dzz);
    edge_candidates(tnew, st, p.pxmy, F, (ax * dxx - ay * dyy) / dxy,
                    (ax * dyy + ay * dxx) / dxy, az, dxy, dzz);
  }
  return ((told > R(0)) ? R(1) : R(-1)) * tnew;
}

}  // namespace fold3d
