// gpu_init.hip — the reference's implicit-surface initializer on the GPU,
// same bits as the host initializer (host_init.c, reference
// Stencil-OpenACC/Fold3dPMM.c:129-389).
//
// Only the sign of IF = ImpSurfF(x, y, z) matters at most nodes (T = sign *
// MAXT); its value matters only at the corners of voxel edges where the surface
// crosses (the crossing point t = h f0 / (f0 - f1)).  So:
//   1. the device evaluates IF everywhere with its own sin/cos; where |IF| is
//      below kAmbiguous (1e-9, far above any libm error on terms of size <= 60)
//      the sign is in doubt;
//   2. every node whose exact value can matter — |IF| below kAmbiguous, an
//      axis neighbour of such a node, or an end of an axis edge whose two signs
//      differ — is listed, and the host evaluates IF there with its own libm,
//      exactly as the host initializer does;
//   3. those exact values (and signs) are written back, and one thread per
//      voxel finds the crossings, the blended points and the 44 node updates
//      of each point, in the host initializer's arithmetic.  A node update
//      keeps the smaller |T| with the node's fixed sign, so concurrent updates
//      are an atomic minimum on the 64-bit pattern (for doubles of one sign the
//      pattern orders by magnitude).
// Every other node keeps sign(IF) * MAXT, whose sign is certain.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/Fold3dPMM.h"
#include "common.h"
#include "solver.h"

namespace fold3d {

constexpr double kAmbiguous = 1e-9;

struct InitGrid {
  int nx, ny, nz;
  long long H0, H1;
  double dx, dy, dz;
  __host__ __device__ long long node(int i, int j, int k) const {
    return i + H0 * (j + H1 * (long long)k);
  }
};

__device__ __forceinline__ double dev_sgn(double v) { return (v > 0) ? 1.0 : -1.0; }

// Fold3dPMM.c:29-33 with the device's sin/cos (value used only for its sign
// unless the host replaces it)
__device__ __forceinline__ double dev_impsurf(double x, double y, double z) {
  return 5.0 + (sin(x) + cos(y + x) + sin(x + y + z)) * 0.4 - z +
         (x - 5.0) * (y - 5.0) * (z - 5) / 55.0;
}

__global__ void init_if_kernel(InitGrid G, long long count, double *__restrict__ IF,
                               double *__restrict__ T, unsigned char *__restrict__ doubt) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < count;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % G.H0), j = (int)((c / G.H0) % G.H1), k = (int)(c / (G.H0 * G.H1));
    const double x = (i - 1) * G.dx + _xmin, y = (j - 1) * G.dy + _ymin, z = (k - 1) * G.dz + _zmin;
    const double f = dev_impsurf(x, y, z);
    IF[c] = f;
    T[c] = dev_sgn(f) * MAXT;
    doubt[c] = fabs(f) < kAmbiguous;
  }
}

// Nodes whose exact IF can matter: pass 0 counts them, pass 1 lists them.
__global__ void init_mark_kernel(InitGrid G, long long count, const double *__restrict__ IF,
                                 const unsigned char *__restrict__ doubt,
                                 unsigned long long *__restrict__ n_exact,
                                 long long *__restrict__ list, long long capacity) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < count;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(c % G.H0), j = (int)((c / G.H0) % G.H1), k = (int)(c / (G.H0 * G.H1));
    const bool pos = IF[c] > 0;
    bool need = doubt[c];
    const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int q = 0; q < 6 && !need; ++q) {
      const int a = i + nb[q][0], b = j + nb[q][1], e = k + nb[q][2];
      if (a < 0 || b < 0 || e < 0 || a >= G.H0 || b >= G.H1 || e > G.nz + 1) continue;
      const long long o = G.node(a, b, e);
      need = doubt[o] || ((IF[o] > 0) != pos);
    }
    if (need) {
      const unsigned long long slot = atomicAdd(n_exact, 1ull);
      if (list && (long long)slot < capacity) list[slot] = c;
    }
  }
}

__global__ void init_patch_kernel(const long long *__restrict__ list, const double *__restrict__ vals,
                                  long long n, double *__restrict__ IF, double *__restrict__ T) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n;
       q += (long long)gridDim.x * blockDim.x) {
    const long long c = list[q];
    IF[c] = vals[q];
    T[c] = dev_sgn(vals[q]) * MAXT;
  }
}

// Fold3dPMM.c:9-26
__device__ __forceinline__ double dev_fold_point(double F, double ax, double ay, double az,
                                                 double x, double y, double z) {
  const double r2 = x * x + y * y + z * z;
  const double ad = ax * x + ay * y + az * z;
  double q = (F * F - (ax * ax + ay * ay + az * az)) * r2 + ad * ad;
  if (q > 0.0) {
    q = ad + sqrt(q);
    if (q > 0.0) return r2 / q;
  }
  return 1.0 * MAXT;
}

// Fold3dPMM.c:35-54 as an atomic minimum of |T| at the node's fixed sign.
__device__ __forceinline__ void dev_init_node(double x, double y, double z, long long c,
                                              double *T) {
  const double t = T[c];  // only its sign is used, and the sign never changes
  const double s = (t < 0.0) ? -1.0 : 1.0;
  const double tn = dev_fold_point(_Fc, s * _axx, s * _ayy, s * _azz, -x, -y, -z);
  if (tn < 1.0 * MAXT) {
    const double v = dev_sgn(t) * tn;
    atomicMin((unsigned long long *)&T[c], (unsigned long long)__double_as_longlong(v));
  }
}

// The 44 node offsets and coordinate shifts of a point in voxel (i,j,k)
// (Fold3dPMM.c:56-112, the same table as the host initializer's).
__constant__ signed char kVoxelDev[44][6] = {
    {0, 0, 0, 0, 0, 0},     {1, 0, 0, 1, 0, 0},     {0, 1, 0, 0, 1, 0},
    {0, 0, 1, 0, 0, 1},     {1, 1, 0, 1, 1, 0},     {1, 0, 1, 1, 0, 1},
    {0, 1, 1, 0, 1, 1},     {1, 1, 1, 1, 1, 1},
    {-1, 0, 0, -1, 0, 0},   {-1, 1, 0, -1, 1, 0},   {-1, 0, 1, -1, 0, 1},
    {-1, 1, 1, -1, 1, 1},   {2, 0, 0, 2, 0, 0},     {2, 1, 0, 2, 1, 0},
    {2, 0, 1, 2, 0, 1},     {2, 1, 1, 2, 1, 1},
    {0, -1, 0, 0, -1, 0},   {1, 0, 0, 1, -1, 0},    {0, -1, 1, 0, -1, 1},
    {1, -1, 1, 1, -1, 1},   {0, 2, 0, 0, 2, 0},     {2, 0, 0, 1, 2, 0},
    {0, 2, 1, 0, 2, 1},     {1, 2, 1, 1, 2, 1},
    {0, 0, -1, 0, 0, -1},   {1, 0, -1, 1, 0, -1},   {0, 1, -1, 0, 1, -1},
    {1, 1, -1, 1, 1, -1},   {0, 0, 2, 0, 0, 2},     {1, 0, 2, 1, 0, 2},
    {0, 1, 2, 0, 1, 2},     {1, 1, 2, 1, 1, 2},
    {0, -1, -1, 0, -1, -1}, {-1, 0, -1, -1, 0, -1}, {-1, -1, 0, -1, -1, 0},
    {0, 2, 2, 0, 2, 2},     {2, 0, 2, 2, 0, 2},     {2, 2, 0, 2, 2, 0},
    {1, -1, -1, 1, -1, -1}, {-1, 1, -1, -1, 1, -1}, {-1, -1, 1, -1, -1, 1},
    {1, 2, 2, 1, 2, 2},     {2, 1, 2, 2, 1, 2},     {2, 2, 1, 2, 2, 1},
};

__device__ __forceinline__ double dev_shifted(double p, int c, double h) {
  return c == 0 ? p : c == 1 ? p - h : c == -1 ? p + h : p - 2 * h;
}

__device__ void dev_update_voxel(const InitGrid &G, double *T, int i, int j, int k, double xp,
                                 double yp, double zp) {
  for (int e = 0; e < 44; ++e) {
    const signed char *r = kVoxelDev[e];
    dev_init_node(dev_shifted(xp, r[3], G.dx), dev_shifted(yp, r[4], G.dy),
                  dev_shifted(zp, r[5], G.dz), G.node(i + r[0], j + r[1], k + r[2]), T);
  }
}

// One voxel: crossings of its 12 edges (Fold3dPMM.c:173-269), then every point
// and its blends (:271-383) in the host initializer's arithmetic.
__global__ void init_voxel_kernel(InitGrid G, const double *__restrict__ IF, double *T) {
  const long long nvx = (long long)(G.nx - 1) * (G.ny - 1) * (G.nz - 1);
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < nvx;
       v += (long long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(v % (G.nx - 1));
    const int j = 1 + (int)((v / (G.nx - 1)) % (G.ny - 1));
    const int k = 1 + (int)(v / ((long long)(G.nx - 1) * (G.ny - 1)));
    double c[8];
    for (int q = 0; q < 8; ++q) c[q] = IF[G.node(i + (q & 1), j + (q >> 1 & 1), k + (q >> 2))];
    // quick reject: no corner sign change and no corner near zero -> no crossing
    bool any_pos = false, any_neg = false, small = false;
    for (int q = 0; q < 8; ++q) {
      any_pos |= c[q] > 0;
      any_neg |= c[q] <= 0;
      small |= fabs(c[q]) < kAmbiguous;
    }
    if (!(any_pos && any_neg) && !small) continue;
    double X[12], Y[12], Z[12];
    int np = 0;
    const double dxx = G.dx, dyy = G.dy, dzz = G.dz;
    const int xe[4] = {0, 4, 2, 6}, ye[4] = {0, 4, 1, 5}, ze[4] = {0, 3, 1, 2};
    for (int e = 0; e < 4; ++e) {
      const double f0 = c[xe[e]], f1 = c[xe[e] + 1];
      const double t = dxx * f0 / (f0 - f1);
      if (0.0 <= t && t <= dxx) {
        X[np] = t;
        Y[np] = (xe[e] & 2) ? dyy : 0.0;
        Z[np] = (xe[e] & 4) ? dzz : 0.0;
        ++np;
      }
    }
    for (int e = 0; e < 4; ++e) {
      const double f0 = c[ye[e]], f1 = c[ye[e] + 2];
      const double t = dyy * f0 / (f0 - f1);
      if (0.0 <= t && t <= dyy) {
        X[np] = (ye[e] & 1) ? dxx : 0.;
        Y[np] = t;
        Z[np] = (ye[e] & 4) ? dzz : 0.0;
        ++np;
      }
    }
    for (int e = 0; e < 4; ++e) {
      const double f0 = c[ze[e]], f1 = c[ze[e] + 4];
      const double t = dzz * f0 / (f0 - f1);
      if (0.0 <= t && t <= dzz) {
        X[np] = (ze[e] & 1) ? dxx : 0.0;
        Y[np] = (ze[e] & 2) ? dyy : 0.0;
        Z[np] = t;
        ++np;
      }
    }
    for (int l = 0; l < np; ++l) {
      dev_update_voxel(G, T, i, j, k, X[l], Y[l], Z[l]);
      for (int back = 1; back <= 5 && back <= l; ++back) {
        const int m = l - back;
        dev_update_voxel(G, T, i, j, k, (2 * X[l] + X[m]) / 3, (2 * Y[l] + Y[m]) / 3,
                         (2 * Z[l] + Z[m]) / 3);
        dev_update_voxel(G, T, i, j, k, (X[l] + 2 * X[m]) / 3, (Y[l] + 2 * Y[m]) / 3,
                         (Z[l] + 2 * Z[m]) / 3);
        if (back == 3) {
          const int b = l - 2, cc = l - 3;
          dev_update_voxel(G, T, i, j, k, (X[l] + X[b] + X[cc]) / 3, (Y[l] + Y[b] + Y[cc]) / 3,
                           (Z[l] + Y[b] + Z[cc]) / 3);
          dev_update_voxel(G, T, i, j, k, (2 * X[l] + X[b] + X[cc]) / 4,
                           (2 * Y[l] + Y[b] + Y[cc]) / 4, (2 * Z[l] + Y[b] + Z[cc]) / 4);
          dev_update_voxel(G, T, i, j, k, (X[l] + 2 * X[b] + X[cc]) / 4,
                           (Y[l] + 2 * Y[b] + Y[cc]) / 4, (Z[l] + 2 * Y[b] + Z[cc]) / 4);
          dev_update_voxel(G, T, i, j, k, (X[l] + X[b] + 2 * X[cc]) / 4,
                           (Y[l] + Y[b] + 2 * Y[cc]) / 4, (Z[l] + Y[b] + 2 * Z[cc]) / 4);
        } else if (back >= 4) {
          const int pairs4[3][2] = {{2, 3}, {2, 4}, {3, 4}};
          const int pairs5[5][2] = {{2, 3}, {2, 4}, {2, 5}, {3, 5}, {3, 4}};
          const int npair = back == 4 ? 3 : 5;
          for (int t = 0; t < npair; ++t) {
            const int b = l - (back == 4 ? pairs4[t][0] : pairs5[t][0]);
            const int cc = l - (back == 4 ? pairs4[t][1] : pairs5[t][1]);
            dev_update_voxel(G, T, i, j, k, (X[l] + X[b] + X[cc]) / 3,
                             (Y[l] + Y[b] + Y[cc]) / 3, (Z[l] + Y[b] + Z[cc]) / 3);
          }
        }
      }
    }
  }
}

int device_initialiser(hipStream_t st, double *T, int nx, int ny, int nz, double dx, double dy,
                       double dz, long long *n_exact_out) {
  InitGrid G;
  G.nx = nx;
  G.ny = ny;
  G.nz = nz;
  G.H0 = nx + 2;
  G.H1 = ny + 2;
  G.dx = dx;
  G.dy = dy;
  G.dz = dz;
  const long long count = G.H0 * G.H1 * (long long)(nz + 2);
  double *IF = nullptr, *vals = nullptr;
  unsigned char *doubt = nullptr;
  unsigned long long *ctr = nullptr;
  long long *list = nullptr;
  int rc = FOLD3D_OK;
  unsigned long long n_exact = 0;
  std::vector<long long> hlist;
  std::vector<double> hvals;
#define INIT_HIP(call)                                                               \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error("fold3d_initialiser_device: %s failed: %s", #call, hipGetErrorString(e_)); \
      rc = e_ == hipErrorOutOfMemory ? FOLD3D_E_NOMEM : FOLD3D_E_HIP;                \
      goto done;                                                                     \
    }                                                                                \
  } while (0)
  INIT_HIP(hipMalloc(&IF, sizeof(double) * count));
  INIT_HIP(hipMalloc(&doubt, count));
  INIT_HIP(hipMalloc(&ctr, sizeof(unsigned long long)));
  INIT_HIP(hipMemsetAsync(ctr, 0, sizeof(unsigned long long), st));
  init_if_kernel<<<4096, 256, 0, st>>>(G, count, IF, T, doubt);
  init_mark_kernel<<<4096, 256, 0, st>>>(G, count, IF, doubt, ctr, nullptr, 0);
  INIT_HIP(hipGetLastError());
  INIT_HIP(hipMemcpyAsync(&n_exact, ctr, sizeof n_exact, hipMemcpyDeviceToHost, st));
  INIT_HIP(hipStreamSynchronize(st));
  if (n_exact > 0) {
    INIT_HIP(hipMalloc(&list, sizeof(long long) * n_exact));
    INIT_HIP(hipMalloc(&vals, sizeof(double) * n_exact));
    INIT_HIP(hipMemsetAsync(ctr, 0, sizeof(unsigned long long), st));
    init_mark_kernel<<<4096, 256, 0, st>>>(G, count, IF, doubt, ctr, list, (long long)n_exact);
    INIT_HIP(hipGetLastError());
    hlist.resize(n_exact);
    hvals.resize(n_exact);
    INIT_HIP(hipMemcpyAsync(hlist.data(), list, sizeof(long long) * n_exact,
                            hipMemcpyDeviceToHost, st));
    INIT_HIP(hipStreamSynchronize(st));
    // the host's libm, exactly as the host initializer evaluates IF
    for (unsigned long long q = 0; q < n_exact; ++q) {
      const long long c = hlist[q];
      const int i = (int)(c % G.H0), j = (int)((c / G.H0) % G.H1), k = (int)(c / (G.H0 * G.H1));
      hvals[q] = ImpSurfF(x(i, dx), y(j, dy), z(k, dz));
    }
    INIT_HIP(hipMemcpyAsync(vals, hvals.data(), sizeof(double) * n_exact, hipMemcpyHostToDevice,
                            st));
    init_patch_kernel<<<1024, 256, 0, st>>>(list, vals, (long long)n_exact, IF, T);
    INIT_HIP(hipGetLastError());
  }
  if (nx >= 2 && ny >= 2 && nz >= 2) {
    init_voxel_kernel<<<4096, 256, 0, st>>>(G, IF, T);
    INIT_HIP(hipGetLastError());
  }
  INIT_HIP(hipStreamSynchronize(st));
done:
#undef INIT_HIP
  (void)hipFree(IF);
  (void)hipFree(doubt);
  (void)hipFree(ctr);
  (void)hipFree(list);
  (void)hipFree(vals);
  if (n_exact_out) *n_exact_out = (long long)n_exact;
  return rc;
}

}  // namespace fold3d
