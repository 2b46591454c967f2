// layout.hip — moves a padded field between axis orders.
//
// The reference stores T[i + H0*(j + H1*k)]: x fastest, z slowest (layout Z,
// Fold3dPMM.h:27-30).  A sweep reads plane p and writes plane p+-1 along its
// march axis; when that axis is the slowest one, a plane is one contiguous
// slab and the three source rows a node needs are neighbours in memory.  The
// engine therefore runs each pair of sweeps in its own order:
//   layout X = (y, z, x)  for the x-sweeps   T[j + H1*(k + H2*i)]
//   layout Z = (x, y, z)  for the z-sweeps   (the reference layout)
//   layout Y = (x, z, y)  for the y-sweeps   T[i + H0*(k + H2*j)]
// and moves the field between them with the kernels below (tiles of the
// (x, y) plane staged through LDS when the fastest axis changes, plain row
// copies when it does not).
#include <hip/hip_runtime.h>

#include <utility>

#include "common.h"
#include "solver.h"

namespace fold3d {

// Transposing moves: a square tile of the (x, y) plane whose rows are 256
// bytes (32 doubles or 64 floats), per block of (tile side) x 8 threads.
template <typename R>
constexpr int kTile = 256 / sizeof(R);
constexpr int kTileRows = 8;
constexpr int kRowLanes = 64;  // other moves: 64 lanes along x
constexpr int kRowRows = 4;    //   by 4 rows of y per block, 16 rows per block

struct Strides3 {
  long long s[3];
};

// Moves that change the fastest axis (x <-> y): the tile is read along the
// input's fastest axis and written along the output's, through LDS.
template <typename R, bool kInFastX>
__global__ __launch_bounds__(kTile<R> *kTileRows) void transpose_kernel(
    const R *__restrict__ in, R *__restrict__ out, Strides3 is, Strides3 os, int Hx, int Hy) {
  constexpr int kT = kTile<R>;
  __shared__ R tile[kT][kT + 1];  // [y - y0][x - x0]
  const int x0 = blockIdx.x * kT, y0 = blockIdx.y * kT;
  const long long z = blockIdx.z;
  const long long ib = z * is.s[2], ob = z * os.s[2];
  const int tx = threadIdx.x;
#pragma unroll
  for (int r = threadIdx.y; r < kT; r += kTileRows) {
    const int x = x0 + (kInFastX ? tx : r), y = y0 + (kInFastX ? r : tx);
    if (x < Hx && y < Hy)
      tile[y - y0][x - x0] = in[ib + x * is.s[0] + y * is.s[1]];
  }
  __syncthreads();
#pragma unroll
  for (int r = threadIdx.y; r < kT; r += kTileRows) {
    const int x = x0 + (kInFastX ? r : tx), y = y0 + (kInFastX ? tx : r);
    if (x < Hx && y < Hy) out[ob + x * os.s[0] + y * os.s[1]] = tile[y - y0][x - x0];
  }
}

// Moves that keep x fastest (layouts Z <-> Y): rows of x are copied as they
// are, 64 lanes per row segment, with no staging.
template <typename R>
__global__ __launch_bounds__(kRowLanes *kRowRows) void restride_kernel(
    const R *__restrict__ in, R *__restrict__ out, Strides3 is, Strides3 os, int Hx, int Hy) {
  constexpr int kRows = 16;  // rows of y per block
  const int x = blockIdx.x * kRowLanes + threadIdx.x, y0 = blockIdx.y * kRows;
  const long long z = blockIdx.z;
  if (x >= Hx) return;
  R v[kRows / kRowRows];
#pragma unroll
  for (int k = 0; k < kRows / kRowRows; ++k) {
    const int y = y0 + threadIdx.y + k * kRowRows;
    if (y < Hy) v[k] = in[z * is.s[2] + x + y * is.s[1]];
  }
#pragma unroll
  for (int k = 0; k < kRows / kRowRows; ++k) {
    const int y = y0 + threadIdx.y + k * kRowRows;
    if (y < Hy) out[z * os.s[2] + x + y * os.s[1]] = v[k];
  }
}

static Strides3 strides_of(const Grid &G) {
  Strides3 s;
  for (int a = 0; a < 3; ++a) s.s[a] = G.stride[a];
  return s;
}

template <typename R>
void permute_layout(hipStream_t st, const R *in, const Grid &Gin, R *out, const Grid &Gout) {
  const int Hx = Gin.n[0] + 2, Hy = Gin.n[1] + 2, Hz = Gin.n[2] + 2;
  // every layout has x or y fastest (stride 1)
  const bool in_x = Gin.stride[0] == 1, out_x = Gout.stride[0] == 1;
  const Strides3 is = strides_of(Gin), os = strides_of(Gout);
  if (in_x == out_x) {
    const dim3 block(kRowLanes, kRowRows);
    if (in_x) {
      const dim3 grid((Hx + kRowLanes - 1) / kRowLanes, (Hy + 15) / 16, Hz);
      restride_kernel<R><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
    } else {  // y fastest in both: swap the roles of x and y
      Strides3 is2 = is, os2 = os;
      std::swap(is2.s[0], is2.s[1]);
      std::swap(os2.s[0], os2.s[1]);
      const dim3 grid((Hy + kRowLanes - 1) / kRowLanes, (Hx + 15) / 16, Hz);
      restride_kernel<R><<<grid, block, 0, st>>>(in, out, is2, os2, Hy, Hx);
    }
  } else {
    constexpr int kT = kTile<R>;
    const dim3 block(kT, kTileRows);
    const dim3 grid((Hx + kT - 1) / kT, (Hy + kT - 1) / kT, Hz);
    if (in_x) transpose_kernel<R, true><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
    else transpose_kernel<R, false><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
  }
}

template void permute_layout<double>(hipStream_t, const double *, const Grid &, double *,
                                     const Grid &);
template void permute_layout<float>(hipStream_t, const float *, const Grid &, float *,
                                    const Grid &);

}  // namespace fold3d
