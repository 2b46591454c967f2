// adjoint.hip — gradient of a receiver traveltime misfit with respect to the
// per-node slowness s = 1/v and the tilt angle theta of the elliptic model.
//
// At the converged forward field every node X is the minimum of its candidates
// over the six sweep directions, so T_X = F_X(T at its winning simplex, s_X,
// N_X).  By the envelope theorem the derivatives of that candidate are
//   dT_X/dT_Y   = barycentric weight of Y at the minimising foot point,
//   dT_X/ds_X   = |w|_N               (w = X - foot point),
//   dT_X/dtheta = s (w^T dN/dtheta w) / (2 |w|_N).
// The adjoint state mu solves mu_X = r_X + sum_Z (dT_Z/dT_X) mu_Z (Z downstream
// of X), a triangular system in the order of T, solved here by reverse-ordered
// plane sweeps (Gauss-Seidel) until mu stops changing.  Before the sweeps each
// node's in-edges are collected once into a 64-bit mask (two bits per
// neighbour direction: which stencil slot of that neighbour names the node),
// so a sweep visits only the neighbours that feed it.  Then
//   dJ/ds_X = mu_X dT_X/ds_X,  dJ/dtheta_X = mu_X dT_X/dtheta_X.
// Nodes no candidate reproduces (the closed-form start around a source) are
// fixed: they pass nothing upstream and contribute no gradient; they are
// counted (held_out) so that a lost gradient path does not go unnoticed.
//
// A node's winning simplex may lie in another sweep direction's upwind plane,
// so an in-edge can join two nodes of the same plane of the sweep being run.
// Those in-plane values are read from a snapshot of mu taken before the
// direction's sweep, never from mu itself while the plane launch writes it:
// every pass computes the same values whatever the order of the waves.
//
// Several sources on one grid and model are solved together: every plane
// launch serves all sources still running (blockIdx.z = source), which gives
// the small planes of one source enough waves to fill the GPU.  Each source's
// values are computed exactly as alone, it stops on its own tolerance, and the
// gradients are added in source order, so the bits are those of one solve per
// source in that order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

#include "common.h"
#include "ellip_local.h"
#include "solver.h"

namespace fold3d {

// Winning stencil of every node (structure of arrays, reference layout).
template <typename R>
struct Stencil {
  int *off0, *off1, *off2;  // linear offsets of the simplex vertices from the node
  R *w0, *w1, *w2;          // their barycentric weights (0 for unused slots)
  R *dTds, *dTdth;          // local partial derivatives
};

template <typename R>
struct AdjArgs {
  int n[3];
  long long stride[3];
  R h[3];
  R speed;
  const R *V;                      // speed field (required for a slowness gradient)
  const R *th, *ph, *ra, *rb;      // per-node metric or null
  R Nc[6];                         // constant inverse metric
  R rel_tol;                       // a candidate within rel_tol of T counts as the winner
};

// dN/dtheta of N = Q diag(1, 1/ra^2, 1/rb^2) Q^T, Q = Rz(phi) Ry(theta).
template <typename R>
__device__ __forceinline__ void dN_dtheta(R theta, R phi, R ra, R rb, R dN[6]) {
  const R ct = cos(theta), st = sin(theta), cp = cos(phi), sp = sin(phi);
  const R Q[3][3] = {{cp * ct, -sp, cp * st}, {sp * ct, cp, sp * st}, {-st, R(0), ct}};
  const R dQ[3][3] = {{-cp * st, R(0), cp * ct}, {-sp * st, R(0), sp * ct}, {-ct, R(0), -st}};
  const R wgt[3] = {R(1), R(1) / (ra * ra), R(1) / (rb * rb)};
  R F[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R acc = R(0);
      for (int k = 0; k < 3; ++k) acc += (dQ[i][k] * Q[j][k] + Q[i][k] * dQ[j][k]) * wgt[k];
      F[i][j] = acc;
    }
  dN[0] = F[0][0];
  dN[1] = F[1][1];
  dN[2] = F[2][2];
  dN[3] = F[0][1];
  dN[4] = F[0][2];
  dN[5] = F[1][2];
}

template <typename R>
__device__ __forceinline__ R quad6(const R M[6], const R w[3]) {
  return M[0] * w[0] * w[0] + M[1] * w[1] * w[1] + M[2] * w[2] * w[2] +
         R(2) * (M[3] * w[0] * w[1] + M[4] * w[0] * w[2] + M[5] * w[1] * w[2]);
}

// Best candidate so far, with what the adjoint needs from it.
template <typename R>
struct Winner {
  R val;
  int nv;                // vertices used (1, 2 or 3)
  long long off[3];      // linear offsets from the node
  R wgt[3];
  R ell;                 // |w|_N
  R wg[3];               // w in global axes
};

template <typename R>
__global__ __launch_bounds__(256) void stencil_kernel(const R *__restrict__ T, AdjArgs<R> A,
                                                      Stencil<R> S,
                                                      unsigned long long *__restrict__ held) {
  const long long total = (long long)A.n[0] * A.n[1] * A.n[2];
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (tid >= total) return;
  const int i = 1 + (int)(tid % A.n[0]);
  const int j = 1 + (int)((tid / A.n[0]) % A.n[1]);
  const int k = 1 + (int)(tid / ((long long)A.n[0] * A.n[1]));
  const long long X = i * A.stride[0] + j * A.stride[1] + k * A.stride[2];
  const R tX = T[X];
  if (!reached(tX)) {  // held fixed: no stencil
    S.off0[X] = S.off1[X] = S.off2[X] = 0;
    S.w0[X] = S.w1[X] = S.w2[X] = R(0);
    S.dTds[X] = S.dTdth[X] = R(0);
    return;
  }
  // Only a candidate <= tX (1 + rel_tol) can make the node a stencil node.  A
  // candidate is its foot point's interpolated value (>= the smallest vertex
  // of the simplex) plus s |w|_N with w from a point of the upwind plane to
  // the node, so at least the smallest vertex plus plane_gap (the least
  // N-distance from the node to that plane, times s, shrunk).  Simplices whose
  // bound exceeds tX (1 + rel_tol) (1 + margin), the margin far above the
  // rounding of a candidate, are skipped: they cannot win, so the stencil is
  // the one the full search finds.
  const R thr = tX * (R(1) + A.rel_tol) * (R(1) + (sizeof(R) == 8 ? R(1e-10) : R(1e-4)));
  const R s = A.V ? R(1) / A.V[X] : R(1) / A.speed;
  R N[6], dN[6];
  if (A.th) {
    tilted_inverse_metric<R>(A.th[X], A.ph[X], A.ra[X], A.rb[X], N);
    dN_dtheta<R>(A.th[X], A.ph[X], A.ra[X], A.rb[X], dN);
  } else {
    for (int q = 0; q < 6; ++q) {
      N[q] = A.Nc[q];
      dN[q] = R(0);
    }
  }
  Winner<R> best;
  best.val = R(1e30);
  best.nv = 0;
  for (int d = 0; d < 6; ++d) {
    const Frame fr = kFrames[d];
    const int cm = (fr.m == 0 ? i : fr.m == 1 ? j : k) - fr.s;  // upwind centre along the march axis
    if (cm < 1 || cm > A.n[fr.m]) continue;
    const int u[3] = {fr.ux, fr.uy, fr.m};
    const R sg[3] = {R(fr.sx), R(fr.sy), R(fr.s)};
    const R hl[3] = {A.h[fr.ux], A.h[fr.uy], A.h[fr.m]};
    const LatticeN<R> L = lattice_of<R>(N, u, sg, hl);
    const long long sm = fr.s * A.stride[fr.m];
    const long long su = fr.sx * A.stride[fr.ux], sv = fr.sy * A.stride[fr.uy];
    const long long c0 = X - sm;
    R p[9];
    long long po[9];
    po[0] = -sm;
    #pragma unroll
    for (int q = 0; q < 8; ++q) po[1 + q] = -sm + kRing[q][0] * su + kRing[q][1] * sv;
    R minv = R(1e30);
    #pragma unroll
    for (int q = 0; q < 9; ++q) {
      p[q] = T[X + po[q]];
      if (reached(p[q])) minv = rmin(p[q], minv);
    }
    (void)c0;
    const R gap = plane_gap(L, s);  // <= s |w|_N of every candidate here
    if (!(minv + gap <= thr)) continue;        // no simplex of this direction can win
    // local lattice coordinates of the patch nodes relative to the target
    auto lat = [&](int q, int &a, int &b) {
      a = q == 0 ? 0 : kRing[q - 1][0];
      b = q == 0 ? 0 : kRing[q - 1][1];
    };
    // (by selects: a local array indexed by the frame's axes would live in
    // scratch memory)
    auto to_global = [&](const R wl[3], R wg[3]) {
      const R g0 = sg[0] * wl[0], g1 = sg[1] * wl[1], g2 = sg[2] * wl[2];
#pragma unroll
      for (int a = 0; a < 3; ++a) wg[a] = u[0] == a ? g0 : u[1] == a ? g1 : g2;
    };
    // vertices
    #pragma unroll
    for (int q = 0; q < 9; ++q) {
      if (!reached(p[q]) || p[q] + gap > thr) continue;
      int a, b;
      lat(q, a, b);
      const R len = sqrt(lat_dot(L, -a, -b, 1, -a, -b, 1));
      const R c = p[q] + s * len;
      if (c < best.val) {
        best.val = c;
        best.nv = 1;
        best.off[0] = po[q];
        best.wgt[0] = R(1);
        best.ell = len;
        const R wl[3] = {-a * hl[0], -b * hl[1], hl[2]};
        to_global(wl, best.wg);
      }
    }
    const R s2 = s * s;
    // edges: spokes 0..7 (centre -> ring q), rims 8..15 (ring q -> ring q+1)
    #pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int q0 = e < 8 ? 0 : 1 + (e - 8), q1 = e < 8 ? 1 + e : 1 + ((e - 8 + 1) % 8);
      const R t0 = p[q0], t1 = p[q1];
      if (!(reached(t0) && reached(t1)) || rmin(t0, t1) + gap > thr) continue;
      int a0, b0, a1, b1;
      lat(q0, a0, b0);
      lat(q1, a1, b1);
      const int ea = a1 - a0, eb = b1 - b0;
      const R g = lat_dot(L, ea, eb, 0, ea, eb, 0);
      const R bb = lat_dot(L, ea, eb, 0, -a0, -b0, 1);
      const R cc = lat_dot(L, -a0, -b0, 1, -a0, -b0, 1);
      const R P = R(1) / g, l0 = P * bb, kap = cc - bb * l0;
      const R dd = t1 - t0, mu = P * dd, r = s2 - dd * mu;
      if (!(r > R(0))) continue;
      const R root = sqrt(kap * r);
      const R lam = l0 - (root / r) * mu;
      if (!(lam >= R(0) && lam <= R(1))) continue;
      const R c = t0 + dd * l0 + root;
      if (c < best.val) {
        best.val = c;
        best.nv = 2;
        best.off[0] = po[q0];
        best.off[1] = po[q1];
        best.wgt[0] = R(1) - lam;
        best.wgt[1] = lam;
        best.ell = s * root / r;
        const R wl[3] = {(-a0 - lam * ea) * hl[0], (-b0 - lam * eb) * hl[1], hl[2]};
        to_global(wl, best.wg);
      }
    }
    // triangles: centre, ring q, ring q+1
    #pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int q1 = 1 + t, q2 = 1 + (t + 1) % 8;
      const R t0 = p[0], t1 = p[q1], t2 = p[q2];
      if (!(reached(t0) && reached(t1) && reached(t2)) || rmin(t0, rmin(t1, t2)) + gap > thr)
        continue;
      int a1, b1, a2, b2;
      lat(q1, a1, b1);
      lat(q2, a2, b2);
      const R g11 = lat_dot(L, a1, b1, 0, a1, b1, 0), g22 = lat_dot(L, a2, b2, 0, a2, b2, 0);
      const R g12 = lat_dot(L, a1, b1, 0, a2, b2, 0);
      const R r1 = lat_dot(L, a1, b1, 0, 0, 0, 1), r2 = lat_dot(L, a2, b2, 0, 0, 0, 1);
      const R rdet = R(1) / (g11 * g22 - g12 * g12);
      const R P11 = g22 * rdet, P12 = -g12 * rdet, P22 = g11 * rdet;
      const R l1 = P11 * r1 + P12 * r2, l2 = P12 * r1 + P22 * r2;
      const R kap = L.C - (r1 * l1 + r2 * l2);
      const R d1 = t1 - t0, d2 = t2 - t0;
      const R m1 = P11 * d1 + P12 * d2, m2 = P12 * d1 + P22 * d2;
      const R r = s2 - (d1 * m1 + d2 * m2);
      if (!(r > R(0))) continue;
      const R root = sqrt(kap * r);
      const R x1 = l1 - (root / r) * m1, x2 = l2 - (root / r) * m2;
      if (!(x1 >= R(0) && x2 >= R(0) && x1 + x2 <= R(1))) continue;
      const R c = t0 + (d1 * l1 + d2 * l2) + root;
      if (c < best.val) {
        best.val = c;
        best.nv = 3;
        best.off[0] = po[0];
        best.off[1] = po[q1];
        best.off[2] = po[q2];
        best.wgt[0] = R(1) - x1 - x2;
        best.wgt[1] = x1;
        best.wgt[2] = x2;
        best.ell = s * root / r;
        const R wl[3] = {(-x1 * a1 - x2 * a2) * hl[0], (-x1 * b1 - x2 * b2) * hl[1], hl[2]};
        to_global(wl, best.wg);
      }
    }
  }
  // a node is a stencil node if its best candidate reproduces its value
  const bool follows = best.nv > 0 && reached(tX) && best.val <= tX * (R(1) + A.rel_tol);
  {  // reached nodes held fixed: one atomic per wave (the lanes still here are reached)
    const unsigned long long hb = __builtin_amdgcn_ballot_w64(!follows);
    if (hb && (threadIdx.x & 63) == __builtin_ctzll(__builtin_amdgcn_ballot_w64(true)))
      atomicAdd(held, (unsigned long long)__builtin_popcountll(hb));
  }
  int o[3] = {0, 0, 0};
  R w[3] = {R(0), R(0), R(0)};
  R gs = R(0), gth = R(0);
  if (follows) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (q < best.nv) {
        o[q] = (int)best.off[q];
        w[q] = best.wgt[q];
      }
    }
    gs = best.ell;
    gth = A.th ? s * quad6(dN, best.wg) / (R(2) * best.ell) : R(0);
  }
  S.off0[X] = o[0];
  S.off1[X] = o[1];
  S.off2[X] = o[2];
  S.w0[X] = w[0];
  S.w1[X] = w[1];
  S.w2[X] = w[2];
  S.dTds[X] = gs;
  S.dTdth[X] = gth;
}

// Neighbour code c = (dx+1) + 3 (dy+1) + 9 (dz+1), skipping c = 13 (the node
// itself): 26 codes.  in_mask[X] holds two bits per code: 0 if the neighbour Z
// in that direction does not list X in its stencil, else 1 + the slot q
// (0..2) at which it does.
__device__ __forceinline__ int code_slot(int c) { return c < 13 ? c : c - 1; }

template <typename R>
__global__ void in_mask_kernel(Stencil<R> S, long long count, long long sx, long long sy,
                               long long sz, unsigned long long *__restrict__ in_mask) {
  for (long long z = blockIdx.x * (long long)blockDim.x + threadIdx.x; z < count;
       z += (long long)gridDim.x * blockDim.x) {
    const int offs[3] = {S.off0[z], S.off1[z], S.off2[z]};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int o = offs[q];
      if (o == 0) continue;
      // X = z + o: Z sits at offset -o from X; decode -o into (dx, dy, dz)
      const long long back = -(long long)o;
      int dzv = 0, dyv = 0, dxv = 0;
      long long rem = back;
      for (int k = -1; k <= 1; ++k)
        for (int j = -1; j <= 1; ++j)
          for (int i = -1; i <= 1; ++i)
            if (i * sx + j * sy + k * sz == rem) {
              dxv = i;
              dyv = j;
              dzv = k;
            }
      const int c = code_slot((dxv + 1) + 3 * (dyv + 1) + 9 * (dzv + 1));
      atomicOr(&in_mask[z + o], (unsigned long long)(q + 1) << (2 * c));
    }
  }
}

// Device view of one source: its forward field, adjoint state, residuals,
// in-edge masks and winning stencils.
template <typename R>
struct AdjSrc {
  const R *T;
  R *mu, *res;
  R *snap;  // mu as it was before the running direction's sweep (same layout)
  unsigned long long *in_mask;
  Stencil<R> S;
};

// snap = mu for every source (blockIdx.y = source), before a direction's sweep.
template <typename R>
__global__ void adjoint_snapshot_kernel(const AdjSrc<R> *__restrict__ src,
                                        const int *__restrict__ done, long long count) {
  const int k = blockIdx.y;
  if (done[k]) return;
  const R *__restrict__ mu = src[k].mu;
  R *__restrict__ snap = src[k].snap;
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < count;
       x += (long long)gridDim.x * blockDim.x)
    snap[x] = mu[x];
}

// Largest change of each source's adjoint values in a pass: kAdjMaxSlots
// words per source, a block updating the one its index picks, so that the
// blocks of a plane do not all meet on one word (the host takes the largest).
constexpr int kAdjMaxSlots = 32;

// One plane of an adjoint sweep for every source still running (source =
// blockIdx.z): mu_X = r_X + sum over the neighbours Z that list X in their
// stencil (in_mask[X]) of weight * mu_Z.
template <typename R>
__global__ __launch_bounds__(256) void adjoint_plane_kernel(const AdjSrc<R> *__restrict__ src,
                                                            const int *__restrict__ done,
                                                            PlaneGeom g, long long sx,
                                                            long long sy, long long sz,
                                                            int march,
                                                            unsigned long long *__restrict__ maxd) {
  const int zsrc = blockIdx.z;
  if (done[zsrc]) return;  // uniform over the block
  R *__restrict__ mu = src[zsrc].mu;
  const R *__restrict__ snap = src[zsrc].snap;
  const R *__restrict__ res = src[zsrc].res;
  const Stencil<R> S = src[zsrc].S;
  const unsigned long long *__restrict__ in_mask = src[zsrc].in_mask;
  const int a = 1 + blockIdx.x * 64 + threadIdx.x;
  const int b = 1 + blockIdx.y * 4 + threadIdx.y;
  R d = R(0);
  if (a <= g.na && b <= g.nb) {
    const long long X = g.base + a * g.Sa + b * g.Sb + g.sm;  // the plane being swept
    unsigned long long m = in_mask[X];
    if (m != 0ull) {
      R acc = res[X];
      // the node's own value, read with the residual rather than after the
      // sum (only this thread writes it; in-plane in-edges read the snapshot)
      const R old = mu[X];
      // decode up to kBatch in-edges first so their loads are in flight together,
      // then add them in mask order (the same order, so the same bits, as one
      // at a time); a node with more in-edges finishes one at a time
      // an in-edge from the plane itself (no offset along the march axis) is
      // read from the snapshot: those values are being rewritten by this launch
      constexpr int kBatch = 6;
      const R *wp[kBatch];
      const R *mp[kBatch];
      long long zs[kBatch];
      int cnt = 0;
      while (m && cnt < kBatch) {
        const int c = __builtin_ctzll(m) >> 1;
        const int q = (int)((m >> (2 * c)) & 3ull) - 1;
        m &= ~(3ull << (2 * c));
        const int cc = c < 13 ? c : c + 1;
        const int dx = cc % 3 - 1, dy = (cc / 3) % 3 - 1, dz = cc / 9 - 1;
        const int along = march == 0 ? dx : march == 1 ? dy : dz;
        zs[cnt] = X + dx * sx + dy * sy + dz * sz;
        wp[cnt] = q == 0 ? S.w0 : q == 1 ? S.w1 : S.w2;
        mp[cnt] = along == 0 ? snap : mu;
        ++cnt;
      }
      R wv[kBatch], mv[kBatch];
#pragma unroll
      for (int k = 0; k < kBatch; ++k)
        if (k < cnt) {
          wv[k] = wp[k][zs[k]];
          mv[k] = mp[k][zs[k]];
        }
#pragma unroll
      for (int k = 0; k < kBatch; ++k)
        if (k < cnt) acc += wv[k] * mv[k];
      while (m) {
        const int c = __builtin_ctzll(m) >> 1;
        const int q = (int)((m >> (2 * c)) & 3ull) - 1;
        m &= ~(3ull << (2 * c));
        const int cc = c < 13 ? c : c + 1;
        const int dx = cc % 3 - 1, dy = (cc / 3) % 3 - 1, dz = cc / 9 - 1;
        const int along = march == 0 ? dx : march == 1 ? dy : dz;
        const long long Z = X + dx * sx + dy * sy + dz * sz;
        const R w = q == 0 ? S.w0[Z] : q == 1 ? S.w1[Z] : S.w2[Z];
        acc += w * (along == 0 ? snap : mu)[Z];
      }
      mu[X] = acc;
      d = fabs(acc - old);
    }
  }
  const double v = wave_max_dt(d);
  if ((threadIdx.x & 63) == 0 && v > 0.0)  // the source's word for this block
    atomicMax(&maxd[zsrc * kAdjMaxSlots + (blockIdx.x + 7u * blockIdx.y) % kAdjMaxSlots],
              (unsigned long long)__double_as_longlong(v));
}

// grad += mu_k * dT/dparam_k over the sources k in order: the same additions,
// in the same order, as one source at a time.
template <typename R>
__global__ void adjoint_gradient_kernel(const AdjSrc<R> *__restrict__ src, int nsrc,
                                        long long count, R *__restrict__ gs,
                                        R *__restrict__ gth) {
  for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < count;
       x += (long long)gridDim.x * blockDim.x) {
    R a = gs[x], t = gth ? gth[x] : R(0);
    bool any = false;
    for (int k = 0; k < nsrc; ++k) {
      const R m = src[k].mu[x];
      if (m != R(0)) {
        a += m * src[k].S.dTds[x];
        if (gth) t += m * src[k].S.dTdth[x];
        any = true;
      }
    }
    if (any) {
      gs[x] = a;
      if (gth) gth[x] = t;
    }
  }
}

template <typename R>
__global__ void scatter_residual_kernel(R *__restrict__ res, const int *__restrict__ idx,
                                        const R *__restrict__ val, int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) res[idx[k]] += val[k];
}

__global__ void adj_zero_u64(unsigned long long *p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0ull;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of one source's scratch per node: in-edge masks (u64), mu, res, three
// weights, two partials (R) and three offsets (int) in the reference layout,
// plus the masks, mu, res and weights again in the x-sweep layout, and the
// snapshot of mu (in the layout of the running direction).
static size_t per_source_bytes(const Grid &G, size_t r) {
  return align256((size_t)G.count * (2 * sizeof(unsigned long long) + 13 * r + 3 * sizeof(int)));
}

template <typename R>
int adjoint_solve(hipStream_t st, int nsrc, const AdjointSource *srcs, const Grid &G,
                  const fold3d_model &m, const double h[3], int max_iters, R *grad_s,
                  R *grad_th, void *scratch, int *iters_out, double *maxd_out,
                  long long *held_out) {
  const long long count = G.count;
  const size_t per = per_source_bytes(G, sizeof(R));
  char *base = (char *)scratch;
  // after the sources: the device table of sources, max|d mu| and done flags
  // x-sweeps run on copies in layout X (x slowest), where a plane x = const is
  // contiguous: in the reference layout its rows are strided by n + 2 and
  // every lane of a wave would load its own cache line
  const Grid GX = G.with_order(kOrderX);
  AdjSrc<R> *d_src = (AdjSrc<R> *)(base + per * nsrc);
  AdjSrc<R> *d_srcX = d_src + nsrc;
  unsigned long long *d_max =
      (unsigned long long *)((char *)d_src + align256(sizeof(AdjSrc<R>) * 2 * nsrc));
  int *d_done = (int *)((char *)d_max + align256(sizeof(unsigned long long) * nsrc * kAdjMaxSlots));
  unsigned long long *d_held = (unsigned long long *)((char *)d_done + align256(sizeof(int) * nsrc));
  std::vector<AdjSrc<R>> hs(2 * nsrc);  // [0, nsrc): reference layout, then layout X
  for (int k = 0; k < nsrc; ++k) {
    char *p = base + per * k;
    AdjSrc<R> &A = hs[k];
    A.T = (const R *)srcs[k].T;
    A.in_mask = (unsigned long long *)p;
    p += count * sizeof(unsigned long long);
    R **fields[7] = {&A.mu, &A.res, &A.S.w0, &A.S.w1, &A.S.w2, &A.S.dTds, &A.S.dTdth};
    for (R **f : fields) {
      *f = (R *)p;
      p += count * sizeof(R);
    }
    A.S.off0 = (int *)p;
    p += count * sizeof(int);
    A.S.off1 = (int *)p;
    p += count * sizeof(int);
    A.S.off2 = (int *)p;
    p += count * sizeof(int);
    AdjSrc<R> &X = hs[nsrc + k];
    memset(&X, 0, sizeof X);
    X.in_mask = (unsigned long long *)p;
    p += count * sizeof(unsigned long long);
    R **fx[5] = {&X.mu, &X.res, &X.S.w0, &X.S.w1, &X.S.w2};
    for (R **f : fx) {
      *f = (R *)p;
      p += count * sizeof(R);
    }
    A.snap = X.snap = (R *)p;
    p += count * sizeof(R);
  }
  F3D_HIP(hipMemsetAsync(scratch, 0, adjoint_scratch_bytes(G, sizeof(R) == 8 ? FOLD3D_F64 : FOLD3D_F32, nsrc), st));
  F3D_HIP(hipMemcpyAsync(d_src, hs.data(), sizeof(AdjSrc<R>) * 2 * nsrc, hipMemcpyHostToDevice, st));
  AdjArgs<R> A;
  memset(&A, 0, sizeof A);
  for (int a = 0; a < 3; ++a) {
    A.n[a] = G.n[a];
    A.stride[a] = G.stride[a];
    A.h[a] = (R)h[a];
  }
  A.speed = (R)m.speed;
  A.V = (const R *)m.speed_field;
  A.th = (const R *)m.theta;
  A.ph = (const R *)m.phi;
  A.ra = (const R *)m.ra;
  A.rb = (const R *)m.rb;
  double N[6];
  sym3_inverse(m.M, N);
  for (int q = 0; q < 6; ++q) A.Nc[q] = (R)N[q];
  A.rel_tol = sizeof(R) == 8 ? (R)1e-12 : (R)1e-6;
  const long long interior = (long long)G.n[0] * G.n[1] * G.n[2];
  for (int k = 0; k < nsrc; ++k) {
    stencil_kernel<R><<<(unsigned)((interior + 255) / 256), 256, 0, st>>>(hs[k].T, A, hs[k].S,
                                                                         d_held + k);
    in_mask_kernel<R><<<2048, 256, 0, st>>>(hs[k].S, count, G.stride[0], G.stride[1],
                                            G.stride[2], hs[k].in_mask);
    const int nrec = srcs[k].nrec;
    if (nrec > 0)
      scatter_residual_kernel<R><<<(nrec + 255) / 256, 256, 0, st>>>(
          hs[k].res, srcs[k].rec_idx, (const R *)srcs[k].rec_res, nrec);
    F3D_HIP(hipGetLastError());
    // nodes no other node lists keep mu = res; the sweeps update the rest
    F3D_HIP(hipMemcpyAsync(hs[k].mu, hs[k].res, count * sizeof(R), hipMemcpyDeviceToDevice, st));
    // the fixed fields of the x-sweeps in layout X (the 8-byte masks move as
    // doubles: the copy keeps their bits)
    const AdjSrc<R> &A = hs[k], &X = hs[nsrc + k];
    permute_layout<double>(st, (const double *)A.in_mask, G, (double *)X.in_mask, GX);
    permute_layout<R>(st, A.res, G, X.res, GX);
    permute_layout<R>(st, A.S.w0, G, X.S.w0, GX);
    permute_layout<R>(st, A.S.w1, G, X.S.w1, GX);
    permute_layout<R>(st, A.S.w2, G, X.S.w2, GX);
    F3D_HIP(hipGetLastError());
  }
  // One sweep pass (all six directions, reverse order, all sources) is captured
  // once in a hipGraph on a stream of its own (the legacy default stream cannot
  // be captured) and replayed; after each pass a source whose adjoint values
  // all changed by no more than its tolerance is marked done and skipped.
  hipStream_t ws = nullptr;
  hipEvent_t ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool capturing = false;
  int rc = FOLD3D_OK;
  int pass = 0, running = nsrc;
  std::vector<unsigned long long> bits((size_t)nsrc * kAdjMaxSlots);
  std::vector<int> done(nsrc, 0);
  for (int k = 0; k < nsrc; ++k) {
    iters_out[k] = 0;
    maxd_out[k] = -1.0;
  }
  const dim3 block(64, 4);
#define ADJ_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      set_error("fold3d_adjoint: %s failed: %s", #call, hipGetErrorString(e_));        \
      rc = e_ == hipErrorOutOfMemory ? FOLD3D_E_NOMEM : FOLD3D_E_HIP;                  \
      goto done;                                                                       \
    }                                                                                  \
  } while (0)
  ADJ_HIP(hipStreamCreateWithFlags(&ws, hipStreamNonBlocking));
  ADJ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  ADJ_HIP(hipEventRecord(ev, st));
  ADJ_HIP(hipStreamWaitEvent(ws, ev, 0));
  ADJ_HIP(hipStreamBeginCapture(ws, hipStreamCaptureModeThreadLocal));
  capturing = true;
  adj_zero_u64<<<(nsrc * kAdjMaxSlots + 255) / 256, 256, 0, ws>>>(d_max, nsrc * kAdjMaxSlots);
  // reverse of the forward order: each direction's planes from last to first;
  // the y- and z-sweeps in the reference layout, the x-sweeps (last) in layout
  // X, with mu moved there and back around them
  for (int dd = 5; dd >= 0; --dd) {
    const Frame &fr = kFrames[dd];
    const bool xs = fr.m == 0;
    const Grid &GL = xs ? GX : G;
    if (dd == 1)
      for (int k = 0; k < nsrc; ++k) permute_layout<R>(ws, hs[k].mu, G, hs[nsrc + k].mu, GX);
    adjoint_snapshot_kernel<R><<<dim3(512, nsrc), 256, 0, ws>>>(xs ? d_srcX : d_src, d_done,
                                                                count);
    const int n = GL.n[fr.m];
    for (int q = 0; q < n; ++q) {
      // sweep plane index along the march axis, against the characteristics
      const int pl = fr.s > 0 ? n - q : 1 + q;
      PlaneGeom g = plane_geom(GL, dd, pl);
      g.sm = 0;  // the plane itself (plane_geom addresses a source plane)
      const dim3 grid((g.na + 63) / 64, (g.nb + 3) / 4, nsrc);
      adjoint_plane_kernel<R><<<grid, block, 0, ws>>>(xs ? d_srcX : d_src, d_done, g,
                                                      GL.stride[0], GL.stride[1], GL.stride[2],
                                                      fr.m, d_max);
    }
    if (dd == 0)
      for (int k = 0; k < nsrc; ++k) permute_layout<R>(ws, hs[nsrc + k].mu, GX, hs[k].mu, G);
  }
  capturing = false;
  ADJ_HIP(hipStreamEndCapture(ws, &graph));
  ADJ_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  for (; pass < max_iters && running > 0;) {
    ADJ_HIP(hipGraphLaunch(exec, ws));
    ++pass;
    ADJ_HIP(hipMemcpyAsync(bits.data(), d_max, sizeof(unsigned long long) * bits.size(),
                           hipMemcpyDeviceToHost, ws));
    ADJ_HIP(hipStreamSynchronize(ws));
    bool changed = false;
    for (int k = 0; k < nsrc; ++k) {
      if (done[k]) continue;
      iters_out[k] = pass;
      unsigned long long m = 0;  // largest of the source's words
      for (int w = 0; w < kAdjMaxSlots; ++w) m = std::max(m, bits[(size_t)k * kAdjMaxSlots + w]);
      memcpy(&maxd_out[k], &m, sizeof(double));
      if (maxd_out[k] <= srcs[k].tol) {
        done[k] = 1;
        --running;
        changed = true;
      }
    }
    if (changed && running > 0) {
      ADJ_HIP(hipMemcpyAsync(d_done, done.data(), sizeof(int) * nsrc, hipMemcpyHostToDevice, ws));
      ADJ_HIP(hipStreamSynchronize(ws));
    }
  }
  adjoint_gradient_kernel<R><<<1024, 256, 0, ws>>>(d_src, nsrc, count, grad_s, grad_th);
  ADJ_HIP(hipGetLastError());
  {
    std::vector<unsigned long long> held(nsrc);
    ADJ_HIP(hipMemcpyAsync(held.data(), d_held, sizeof(unsigned long long) * nsrc,
                           hipMemcpyDeviceToHost, ws));
    ADJ_HIP(hipStreamSynchronize(ws));
    for (int k = 0; k < nsrc; ++k) held_out[k] = (long long)held[k];
  }
  // later work on the caller's stream sees the gradients
  ADJ_HIP(hipEventRecord(ev, ws));
  ADJ_HIP(hipStreamWaitEvent(st, ev, 0));
done:
#undef ADJ_HIP
  if (capturing) {
    hipGraph_t g2 = nullptr;
    (void)hipStreamEndCapture(ws, &g2);
    if (g2) (void)hipGraphDestroy(g2);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ws) {
    (void)hipStreamSynchronize(ws);
    (void)hipStreamDestroy(ws);
  }
  if (ev) (void)hipEventDestroy(ev);
  return rc;
}

size_t adjoint_scratch_bytes(const Grid &G, int precision, int nsrc) {
  const size_t r = precision == FOLD3D_F64 ? 8 : 4;
  const size_t tab = sizeof(AdjSrc<double>) > sizeof(AdjSrc<float>) ? sizeof(AdjSrc<double>)
                                                                    : sizeof(AdjSrc<float>);
  return per_source_bytes(G, r) * nsrc + align256(tab * 2 * nsrc) +
         align256(sizeof(unsigned long long) * nsrc * kAdjMaxSlots) + align256(sizeof(int) * nsrc) +
         align256(sizeof(unsigned long long) * nsrc);
}

template int adjoint_solve<double>(hipStream_t, int, const AdjointSource *, const Grid &,
                                   const fold3d_model &, const double *, int, double *, double *,
                                   void *, int *, double *, long long *);
template int adjoint_solve<float>(hipStream_t, int, const AdjointSource *, const Grid &,
                                  const fold3d_model &, const double *, int, float *, float *,
                                  void *, int *, double *, long long *);

}  // namespace fold3d
