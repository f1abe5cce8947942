// implicit.hip -- the marker coupling matrix of implicit IB forcing.
//
// The reference's implicit direct forcing (src/pyib/solver/ib_immersed.py:122-174) solves
// M f = (u_desired - interp(vc)) / dt with M = D S: D the interpolation weights (w hh^3) and S
// the spreading weights (w ds^3) of the marker stencils, so M[i][j] = hh^3 sum over the cells c
// shared by markers i and j of w_i(c) w_j(c) ds_j^3.  Each marker's row is formed by one thread:
// for each of its taps it finds the tap's cell in the cell-sorted tap list of the spreading
// (pyib_ib_sort) and adds the products with every tap in that cell's run, in sorted order.  Rows
// are disjoint, so the matrix is written without atomics and is bitwise reproducible.
#include "common.h"

namespace pyib {
namespace {

__global__ void k_ib_coupling(int64_t nmark, const int32_t* __restrict__ cell, const double* __restrict__ w,
                              const int32_t* __restrict__ keys, const int32_t* __restrict__ vals,
                              const double* __restrict__ ds3, double hh3, double* __restrict__ M) {
  const int64_t ntap = nmark * 27;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nmark; i += (int64_t)gridDim.x * blockDim.x) {
    double* row = M + i * nmark;
    for (int t = 0; t < 27; ++t) {
      const int c = cell[27 * i + t];
      const double wt = w[27 * i + t];
      if (c < 0 || wt == 0.0) continue;
      int64_t lo = 0, hi = ntap;          // first sorted tap with key >= c
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < c) lo = mid + 1; else hi = mid;
      }
      for (int64_t u = lo; u < ntap && keys[u] == c; ++u) {
        const int v = vals[u];
        const int64_t j = v / 27;
        row[j] += hh3 * wt * (w[v] * ds3[j]);
      }
    }
  }
}

__global__ void k_fill(int64_t n, double v, double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

}  // namespace
}  // namespace pyib

extern "C" {
using namespace pyib;

int pyib_ib_coupling(int64_t nmark, const int32_t* cell, const double* w, const int32_t* keys_sorted,
                     const int32_t* vals_sorted, const double* ds3, double hh3, double* M, void* stream) {
  if (nmark < 0 || nmark > 65536) return fail_msg("pyib_ib_coupling: marker count outside [0, 65536]");
  auto st = S(stream);
  k_fill<<<grid_for(nmark * nmark), kBlock, 0, st>>>(nmark * nmark, 0.0, M);
  k_ib_coupling<<<grid_for(nmark), kBlock, 0, st>>>(nmark, cell, w, keys_sorted, vals_sorted, ds3, hh3, M);
  return check_launch("pyib_ib_coupling");
}

}  // extern "C"
