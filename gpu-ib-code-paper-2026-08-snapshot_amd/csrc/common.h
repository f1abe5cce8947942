// common.h -- device helpers shared by the gfx950 kernels of libpyib_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdint>

#include "../../include/pyib_hip.h"

namespace pyib {

constexpr int kBlock = 256;
constexpr int kMaxGrid = 1 << 20;
constexpr int kRedGrid = PYIB_MAX_PARTIALS;   // reduction blocks
constexpr int kPerThread = 4;                 // elements per thread in the reducing kernels
constexpr int kFinBlock = 1024;               // threads of the one-block finalize kernels
enum { S_ETA = 0, S_ETA_NEW = 1, S_DEN = 2, S_ALPHA = 3, S_BETA = 4, S_RR = 5, S_ERR = 6, S_L2I = 7 };

int fail(const char* where, hipError_t e);
int fail_msg(const char* msg);
inline int check_launch(const char* where) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(where, e);
}
inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  return static_cast<int>(g > kMaxGrid ? kMaxGrid : g);
}
inline int red_grid_for(int64_t n) {
  int64_t g = (n + kBlock * kPerThread - 1) / (kBlock * kPerThread);
  if (g < 1) g = 1;
  return static_cast<int>(g > kRedGrid ? kRedGrid : g);
}

template <typename T>
__device__ __forceinline__ T sload(const T* __restrict__ x, int v) {
  return v >= 0 ? x[v] : -x[-v - 1];
}

// Block-wide sum of one double (fixed tree => deterministic); valid in thread 0.
// sh holds blockDim.x/64 <= kFinBlock/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

// max that propagates NaN (NumPy's max does), so non-finite diagnostics are not hidden
__device__ __forceinline__ double nanmax(double a, double b) { return (a != a || b != b) ? (a + b) : fmax(a, b); }

__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = nanmax(r, sh[i]);
  }
  __syncthreads();
  return r;
}

// Sum nsum rows of `nb` partials each: out[k] = sum_b partials[k*kRedGrid + b].
__global__ void k_finalize(const double* __restrict__ partials, int nb, int nsum, double* __restrict__ out);

}  // namespace pyib
