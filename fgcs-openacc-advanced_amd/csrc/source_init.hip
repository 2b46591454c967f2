// source_init.hip — the start field of an elliptic-model solve from a point
// source: T = sqrt(d^T N d) / v at the interior nodes within `radius` cells
// (max norm) of the source, N = M^-1 of the metric at the source, MAXT
// elsewhere and on the ghost layer.  Node (i,j,k) sits at ((i-1)dx, (j-1)dy,
// (k-1)dz).  The same rule the workloads of the tests and benchmarks use.
#include <hip/hip_runtime.h>

#include "../../include/Fold3dPMM.h"
#include "common.h"
#include "solver.h"

namespace fold3d {

template <typename R>
__global__ void point_source_kernel(R *__restrict__ T, int nx, int ny, int nz, double dx,
                                    double dy, double dz, double sx, double sy, double sz,
                                    double n0, double n1, double n2, double n3, double n4,
                                    double n5, double v, double r) {
  const long long H0 = nx + 2, H1 = ny + 2, count = H0 * H1 * (long long)(nz + 2);
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < count;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % H0), j = (int)((c / H0) % H1), k = (int)(c / (H0 * H1));
    const double ex = (i - 1) * dx - sx, ey = (j - 1) * dy - sy, ez = (k - 1) * dz - sz;
    const bool inside = i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz;
    const bool near = fabs(ex) <= r * dx && fabs(ey) <= r * dy && fabs(ez) <= r * dz;
    double t = 1.0 * MAXT;
    if (inside && near) {
      // Left to right, as the host start field of the tests sums it.
      double q = n0 * ex * ex;
      q = q + n1 * ey * ey;
      q = q + n2 * ez * ez;
      q = q + 2.0 * n3 * ex * ey;
      q = q + 2.0 * n4 * ex * ez;
      q = q + 2.0 * n5 * ey * ez;
      t = sqrt(q) / v;
    }
    T[c] = (R)t;
  }
}

int point_source_start(hipStream_t st, void *T, int precision, int nx, int ny, int nz,
                       const double h[3], const double src[3], const double N[6], double v,
                       double radius) {
  if (precision == FOLD3D_F64)
    point_source_kernel<double><<<2048, 256, 0, st>>>(
        (double *)T, nx, ny, nz, h[0], h[1], h[2], src[0], src[1], src[2], N[0], N[1], N[2],
        N[3], N[4], N[5], v, radius);
  else
    point_source_kernel<float><<<2048, 256, 0, st>>>(
        (float *)T, nx, ny, nz, h[0], h[1], h[2], src[0], src[1], src[2], N[0], N[1], N[2],
        N[3], N[4], N[5], v, radius);
  F3D_HIP(hipGetLastError());
  return FOLD3D_OK;
}

}  // namespace fold3d
