// common.h — host-side plumbing shared by the library's translation units:
// status codes, the thread-local error message, HIP call checking, the sweep
// direction frames and the padded grid geometry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/fold3d.h"

namespace fold3d {

void set_error(const char *fmt, ...);

// A failed call is reported and its error taken off the runtime's last-error
// slot, so that a later hipGetLastError() check (after a launch) does not
// report it a second time: a failed allocation leaves the handle usable.
#define F3D_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::fold3d::set_error("%s failed: %s (%s:%d)", #call,                   \
                          hipGetErrorString(e_), __FILE__, __LINE__);       \
      (void)hipGetLastError();                                              \
      return e_ == hipErrorOutOfMemory ? FOLD3D_E_NOMEM : FOLD3D_E_HIP;     \
    }                                                                       \
  } while (0)

#define F3D_CHECK(cond, code, ...)        \
  do {                                    \
    if (!(cond)) {                        \
      ::fold3d::set_error(__VA_ARGS__);   \
      return (code);                      \
    }                                     \
  } while (0)

// Sweep frames.  Each direction marches along axis m in steps of s (+1/-1);
// the upwind plane's local x^ is axis ux with sign sx, local y^ is uy with
// sign sy.  Read off the loads and model constants of the reference kernels
// (AN_ISO_1_OpenACC.c:195-212, 348-364, 500-516, 652-668, 805-821, 960-976)
// and the spacing arguments of its driver (:1198-1235).
struct Frame {
  int m, s, ux, sx, uy, sy;
};
constexpr Frame kFrames[6] = {
    {0, +1, 2, +1, 1, +1},  // x_down: i -> i+1, x^ = +z, y^ = +y
    {0, -1, 2, -1, 1, -1},  // x_up:   i -> i-1, x^ = -z, y^ = -y
    {2, +1, 1, +1, 0, +1},  // k_up:   k -> k+1, x^ = +y, y^ = +x
    {2, -1, 1, -1, 0, -1},  // k_down: k -> k-1, x^ = -y, y^ = -x
    {1, -1, 0, -1, 2, -1},  // y_up:   j -> j-1, x^ = -x, y^ = -z
    {1, +1, 0, +1, 2, +1},  // y_down: j -> j+1, x^ = +x, y^ = +z
};

// Axis orders (fastest to slowest) of the layouts the engine uses.  kOrderZ
// is the reference's T[i + H0*(j + H1*k)]; kOrderX / kOrderY put x / y
// slowest so an x- / y-sweep plane is one contiguous slab.
constexpr int kOrderZ[3] = {0, 1, 2};
constexpr int kOrderX[3] = {1, 2, 0};
constexpr int kOrderY[3] = {0, 2, 1};

// The largest |dT| of an iteration (tolerance stop), kept as the bit pattern
// of a non-negative double, which orders like its value, in kMaxSlots words.
// A block updates the word its block index picks: the thousands of blocks of
// a plane launch then share each word with a few dozen others rather than
// all meeting on one, whose serialised atomics cost more than the update.
constexpr int kMaxSlots = 256;
// Largest |dT| of a whole wave (every lane's d >= 0; inactive lanes pass 0).
// A wave in which no lane changed skips the shuffles: near convergence most
// waves change nothing.
__device__ __forceinline__ double wave_max_dt(double d) {
  if (__builtin_amdgcn_ballot_w64(d > 0.0) == 0) return 0.0;
  double v = d;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// FP32 fields: the maximum taken in float and widened once (widening is exact
// and monotone, so the result is the one of the double form).
__device__ __forceinline__ double wave_max_dt(float d) {
  if (__builtin_amdgcn_ballot_w64(d > 0.0f) == 0) return 0.0;
  float v = d;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return (double)v;
}
__device__ __forceinline__ void max_dt_update(unsigned long long *slots, double v) {
  const unsigned b = blockIdx.x + 7u * blockIdx.y + 13u * blockIdx.z;
  atomicMax(&slots[b % kMaxSlots], (unsigned long long)__double_as_longlong(v));
}
// Clears the words: one block of kMaxSlots threads.
static __global__ void max_dt_reset_kernel(unsigned long long *slots) { slots[threadIdx.x] = 0ull; }
// Largest word, on the host, once the stream's work so far has finished.
static inline hipError_t max_dt_read(hipStream_t st, const unsigned long long *slots,
                                     double *out) {
  unsigned long long h[kMaxSlots];
  hipError_t e = hipMemcpyAsync(h, slots, sizeof h, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  unsigned long long m = 0;
  for (int k = 0; k < kMaxSlots; ++k) m = h[k] > m ? h[k] : m;
  memcpy(out, &m, sizeof m);
  return e;
}

// The source rows b - 1, b, b + 1 of a plane step (66 values each, from
// lane-axis index a0) staged through registers into LDS by kThreads threads.
// load() only issues the loads, so the caller can issue its other loads of
// the step (target value, speed, metric) before anything waits: every thread
// loads, from an address inside the padded grid ((a0, b - 1) always is), and
// the threads outside the rows or the plane drop what they read.  (A load
// under a branch lets the compiler wait for it inside the branch, before the
// next load is issued.)
template <typename R, int kThreads>
struct RowStage {
  static constexpr int kStage = 3 * 66, kPer = (kStage + kThreads - 1) / kThreads;
  R v[kPer];
  bool ok[kPer];
  __device__ __forceinline__ void load(const R *__restrict__ T, long long base, long long Sa,
                                       long long Sb, int a0, int b, int na, int thread) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int idx = thread + q * kThreads;
      const int la = idx % 66, lb = idx / 66;
      ok[q] = idx < kStage && a0 + la <= na + 1;
      v[q] = T[base + (long long)(ok[q] ? a0 + la : a0) * Sa +
               (long long)(b - 1 + (ok[q] ? lb : 0)) * Sb];
    }
  }
  __device__ __forceinline__ void store(R (&s)[3][66], int thread) const {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int idx = thread + q * kThreads;
      if (ok[q]) s[idx / 66][idx % 66] = v[q];
    }
  }
};

// Padded grid: nodes 1..n[a] on each axis, ghosts at 0 and n[a]+1.
struct Grid {
  int n[3];
  long long stride[3];
  long long count;  // (nx+2)(ny+2)(nz+2)
  Grid(int nx, int ny, int nz, const int order[3] = kOrderZ) {
    n[0] = nx;
    n[1] = ny;
    n[2] = nz;
    long long s = 1;
    for (int q = 0; q < 3; ++q) {
      stride[order[q]] = s;
      s *= (long long)n[order[q]] + 2;
    }
    count = s;
  }
  Grid with_order(const int order[3]) const { return Grid(n[0], n[1], n[2], order); }
  // Planes of a sweep along axis m: forward 1..n-1, backward n..2.
  int planes(int m) const { return n[m] - 1; }
};

}  // namespace fold3d
