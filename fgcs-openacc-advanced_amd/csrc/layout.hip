// layout.hip — reorders a padded field between the reference layout
//   A: T[i + H0*(j + H1*k)]          (i fastest; Fold3dPMM.h:27-30)
// and the x-march layout
//   B: T[j + H1*(k + H2*i)]          (j fastest, i slowest)
// in which a plane of constant i is one contiguous slab whose rows run along
// j, so the x-sweeps read and write whole rows like the y- and z-sweeps do in A.
// Both directions stage a 32x32 tile through LDS so reads and writes are
// coalesced.
#include <hip/hip_runtime.h>

#include "common.h"
#include "solver.h"

namespace fold3d {

constexpr int kTile = 32;
constexpr int kTileRows = 8;

// A -> B: for each k, B[i][k][j] = A[k][j][i].
template <typename R>
__global__ __launch_bounds__(kTile *kTileRows) void to_xmarch_kernel(
    const R *__restrict__ A, R *__restrict__ B, int H0, int H1, int H2) {
  __shared__ R tile[kTile][kTile + 1];
  const int i0 = blockIdx.x * kTile, j0 = blockIdx.y * kTile, k = blockIdx.z;
  const long long plane = (long long)H0 * H1 * k;
#pragma unroll
  for (int r = threadIdx.y; r < kTile; r += kTileRows) {
    const int i = i0 + threadIdx.x, j = j0 + r;
    if (i < H0 && j < H1) tile[r][threadIdx.x] = A[plane + (long long)H0 * j + i];
  }
  __syncthreads();
#pragma unroll
  for (int r = threadIdx.y; r < kTile; r += kTileRows) {
    const int j = j0 + threadIdx.x, i = i0 + r;
    if (i < H0 && j < H1) B[(long long)H1 * ((long long)H2 * i + k) + j] = tile[threadIdx.x][r];
  }
}

// B -> A: for each k, A[k][j][i] = B[i][k][j].
template <typename R>
__global__ __launch_bounds__(kTile *kTileRows) void from_xmarch_kernel(
    const R *__restrict__ B, R *__restrict__ A, int H0, int H1, int H2) {
  __shared__ R tile[kTile][kTile + 1];
  const int j0 = blockIdx.x * kTile, i0 = blockIdx.y * kTile, k = blockIdx.z;
#pragma unroll
  for (int r = threadIdx.y; r < kTile; r += kTileRows) {
    const int j = j0 + threadIdx.x, i = i0 + r;
    if (i < H0 && j < H1) tile[r][threadIdx.x] = B[(long long)H1 * ((long long)H2 * i + k) + j];
  }
  __syncthreads();
  const long long plane = (long long)H0 * H1 * k;
#pragma unroll
  for (int r = threadIdx.y; r < kTile; r += kTileRows) {
    const int i = i0 + threadIdx.x, j = j0 + r;
    if (i < H0 && j < H1) A[plane + (long long)H0 * j + i] = tile[threadIdx.x][r];
  }
}

template <typename R>
void to_xmarch(hipStream_t st, const R *A, R *B, const Grid &G) {
  const int H0 = G.n[0] + 2, H1 = G.n[1] + 2, H2 = G.n[2] + 2;
  dim3 block(kTile, kTileRows);
  dim3 grid((H0 + kTile - 1) / kTile, (H1 + kTile - 1) / kTile, H2);
  to_xmarch_kernel<R><<<grid, block, 0, st>>>(A, B, H0, H1, H2);
}

template <typename R>
void from_xmarch(hipStream_t st, const R *B, R *A, const Grid &G) {
  const int H0 = G.n[0] + 2, H1 = G.n[1] + 2, H2 = G.n[2] + 2;
  dim3 block(kTile, kTileRows);
  dim3 grid((H1 + kTile - 1) / kTile, (H0 + kTile - 1) / kTile, H2);
  from_xmarch_kernel<R><<<grid, block, 0, st>>>(B, A, H0, H1, H2);
}

template void to_xmarch<double>(hipStream_t, const double *, double *, const Grid &);
template void to_xmarch<float>(hipStream_t, const float *, float *, const Grid &);
template void from_xmarch<double>(hipStream_t, const double *, double *, const Grid &);
template void from_xmarch<float>(hipStream_t, const float *, float *, const Grid &);

}  // namespace fold3d
