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
// and moves the field between them with the kernel below (a 32x32 tile of the
// (x, y) plane per block, staged through LDS when the fastest axis changes).
#include <hip/hip_runtime.h>

#include "common.h"
#include "solver.h"

namespace fold3d {

constexpr int kTile = 32;
constexpr int kTileRows = 8;

struct Strides3 {
  long long s[3];
};

template <typename R, bool kInFastX, bool kOutFastX>
__global__ __launch_bounds__(kTile *kTileRows) void permute_kernel(
    const R *__restrict__ in, R *__restrict__ out, Strides3 is, Strides3 os, int Hx, int Hy) {
  __shared__ R tile[kTile][kTile + 1];  // [y - y0][x - x0]
  const int x0 = blockIdx.x * kTile, y0 = blockIdx.y * kTile;
  const long long z = blockIdx.z;
  const long long ib = z * is.s[2], ob = z * os.s[2];
  const int tx = threadIdx.x;
#pragma unroll
  for (int r = threadIdx.y; r < kTile; r += kTileRows) {
    const int x = x0 + (kInFastX ? tx : r), y = y0 + (kInFastX ? r : tx);
    if (x < Hx && y < Hy)
      tile[y - y0][x - x0] = in[ib + x * is.s[0] + y * is.s[1]];
  }
  __syncthreads();
#pragma unroll
  for (int r = threadIdx.y; r < kTile; r += kTileRows) {
    const int x = x0 + (kOutFastX ? tx : r), y = y0 + (kOutFastX ? r : tx);
    if (x < Hx && y < Hy) out[ob + x * os.s[0] + y * os.s[1]] = tile[y - y0][x - x0];
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
  const bool in_x = Gin.stride[0] == 1, out_x = Gout.stride[0] == 1;
  const dim3 block(kTile, kTileRows);
  const dim3 grid((Hx + kTile - 1) / kTile, (Hy + kTile - 1) / kTile, Hz);
  const Strides3 is = strides_of(Gin), os = strides_of(Gout);
  if (in_x && out_x) permute_kernel<R, true, true><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
  else if (in_x) permute_kernel<R, true, false><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
  else if (out_x) permute_kernel<R, false, true><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
  else permute_kernel<R, false, false><<<grid, block, 0, st>>>(in, out, is, os, Hx, Hy);
}

template void permute_layout<double>(hipStream_t, const double *, const Grid &, double *,
                                     const Grid &);
template void permute_layout<float>(hipStream_t, const float *, const Grid &, float *,
                                    const Grid &);

}  // namespace fold3d
