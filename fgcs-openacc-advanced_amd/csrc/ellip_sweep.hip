// ellip_sweep.hip — plane-step kernels and sweep driver of the tilted
// elliptic model (constant or per-node metric, constant or per-node speed),
// FP64 and FP32.  Same plane order and frames as the reference's fold-model
// driver (AN_ISO_1_OpenACC.c:1192-1237); convergence by max|dT| < tol.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "common.h"
#include "ellip_local.h"
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
      EllipConsts<R> K;
      ellip_constants<R>(N, A.u, A.sg, A.hl, K);
      tn = ellip_node_update<R>(told, p, s, K);
    } else {
      tn = ellip_node_update<R>(told, p, s, A.K);
    }
    if (tn != told) T[t] = tn;
    if (kReduce) d = told - tn;
  }
  if (kReduce) {
    double v = (double)d;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0 && v > 0.0)
      atomicMax(max_dt, (unsigned long long)__double_as_longlong(v));
  }
}

__global__ void ellip_zero_u64(unsigned long long *p) { *p = 0ull; }

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
    if (tol > 0) ellip_zero_u64<<<1, 1, 0, st>>>(d_max);
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
      unsigned long long bits = 0;
      F3D_HIP(hipMemcpyAsync(&bits, d_max, sizeof bits, hipMemcpyDeviceToHost, st));
      F3D_HIP(hipStreamSynchronize(st));
      memcpy(&last, &bits, sizeof last);
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

}  // namespace fold3d
