// metric_prep.h — the per-node metric of the elliptic model, written once
// per solve into the two layouts the sweeps read it in.
//
// For each node: the slowness 1/v and the six products P_ab = (N_ab h_a) h_b
// of N = M^-1 (from the tilt and axis-ratio fields, or the model's constant
// metric) with the spacings (lattice_products), seven fields in all.  The
// x-sweeps read them in the x-sweep layout (y fastest, then z, then x), the
// z- and y-sweeps in the y-sweep layout (x fastest, then z, then y).  The inputs are in the reference layout (x, y, z).  A block takes one
// 32x32 tile of one (x, y) plane, evaluates each of its nodes once, writes the
// tile into the y-sweep layout directly (x stays fastest) and, through LDS,
// into the x-sweep layout.  The plane z is the fastest block index, so the
// blocks that run together write neighbouring rows of both layouts (a tile's
// own 32 rows lie a whole plane apart in them).
#pragma once

#include <hip/hip_runtime.h>

#include "ellip_local.h"  // tilted_inverse_metric, lattice_products

namespace fold3d {

constexpr int kPrepTile = 32, kPrepRows = 8;

struct MetricPrep {
  long long count;               // nodes of one field (padded grid)
  int hx, hy, hz;                // padded extents
  long long sz[3], sx[3], sy[3];  // strides: reference, x-sweep and y-sweep layouts
};

// grid (hz, ceil(hx / 32), ceil(hy / 32)), block (32, 8)
template <typename R>
__global__ __launch_bounds__(kPrepTile *kPrepRows) void metric_prep_kernel(
    MetricPrep g, const R *__restrict__ V, R speed, const R *__restrict__ th,
    const R *__restrict__ ph, const R *__restrict__ ra, const R *__restrict__ rb, R Nc0, R Nc1,
    R Nc2, R Nc3, R Nc4, R Nc5, R hx, R hy, R hz, R *__restrict__ MX, R *__restrict__ MY) {
  constexpr int kQ = kPrepTile / kPrepRows;  // nodes per thread
  __shared__ R tile[kPrepTile][kPrepTile + 1];  // [y - y0][x - x0]
  const long long z = blockIdx.x;
  const int x0 = blockIdx.y * kPrepTile, y0 = blockIdx.z * kPrepTile;
  const int tx = threadIdx.x, ty = threadIdx.y;
  // Every input of the thread's kQ nodes is loaded before any is used (from
  // addresses clamped into the grid; a node outside it drops its values):
  // loads under the node's bounds test were each waited for before the next
  // node's were issued.
  R tv[kQ], pv[kQ], av[kQ], bv[kQ], vv[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int x = min(x0 + tx, g.hx - 1), y = min(y0 + ty + q * kPrepRows, g.hy - 1);
    const long long iz = x * g.sz[0] + y * g.sz[1] + z * g.sz[2];
    if (th) {  // (uniform)
      tv[q] = th[iz];
      pv[q] = ph[iz];
      av[q] = ra[iz];
      bv[q] = rb[iz];
    }
    vv[q] = V ? V[iz] : speed;
  }
  R val[kQ][7];
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int x = x0 + tx, y = y0 + ty + q * kPrepRows;
    R N[6] = {Nc0, Nc1, Nc2, Nc3, Nc4, Nc5};
    if (th) tilted_inverse_metric<R>(tv[q], pv[q], av[q], bv[q], N);
    val[q][0] = R(1) / vv[q];
    const R h[3] = {hx, hy, hz};
    lattice_products<R>(N, h, &val[q][1]);
    if (x < g.hx && y < g.hy) {
      const long long iy = x * g.sy[0] + y * g.sy[1] + z * g.sy[2];
#pragma unroll
      for (int f = 0; f < 7; ++f) MY[f * g.count + iy] = val[q][f];
    }
  }
#pragma unroll
  for (int f = 0; f < 7; ++f) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) tile[ty + q * kPrepRows][tx] = val[q][f];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const int x = x0 + ty + q * kPrepRows, y = y0 + tx;
      if (x < g.hx && y < g.hy)
        MX[f * g.count + x * g.sx[0] + y * g.sx[1] + z * g.sx[2]] = tile[tx][ty + q * kPrepRows];
    }
    __syncthreads();
  }
}

// Launch for one grid: `ref`, `lx` and `ly` are the strides of the reference,
// x-sweep and y-sweep layouts; N holds the constant metric's inverse (used
// when th is null); h the grid spacings.
template <typename R>
inline void launch_metric_prep(hipStream_t st, long long count, const int n[3],
                               const long long ref[3], const long long lx[3],
                               const long long ly[3], const R *V, R speed, const R *th,
                               const R *ph, const R *ra, const R *rb, const double N[6],
                               const double h[3], R *MX, R *MY) {
  MetricPrep g;
  g.count = count;
  g.hx = n[0] + 2;
  g.hy = n[1] + 2;
  g.hz = n[2] + 2;
  for (int a = 0; a < 3; ++a) {
    g.sz[a] = ref[a];
    g.sx[a] = lx[a];
    g.sy[a] = ly[a];
  }
  const dim3 grid(g.hz, (g.hx + kPrepTile - 1) / kPrepTile, (g.hy + kPrepTile - 1) / kPrepTile);
  metric_prep_kernel<R><<<grid, dim3(kPrepTile, kPrepRows), 0, st>>>(
      g, V, speed, th, ph, ra, rb, (R)N[0], (R)N[1], (R)N[2], (R)N[3], (R)N[4], (R)N[5], (R)h[0],
      (R)h[1], (R)h[2], MX, MY);
}

}  // namespace fold3d
