// grid.hip -- structured per-level kernels of the elliptic operator (gfx950).
//
// The CG of Driver.helm_solve (ib_driver_np.py:311-410) runs on the internal
// layout: per level, dense x/y/z edge blocks and a cell block, +axis oriented.
// A p = rot(integr(V(interp(curl p)))) (ib_core_np.py:213-220) is applied in
// three passes; each pass computes the rows flagged regular with the
// constant-coefficient stencil of its level (one thread per row, the fastest
// index k on consecutive lanes so every stencil read is a coalesced stream)
// and the remaining rows from explicit composite rows (level interfaces,
// outer boundary).  Cell vectors are stored component-major (SoA).
//
//   G1  vc_d(c)  = (U_d(lo) + U_d(hi)) / (2h^2)       U = discrete curl of p
//   V   y(c)     = rou x - (hdt/h^3)(a - b),  a = sum_up (h visc)(x_n - x), b = sum_dn (h visc)(x - x_n)
//   G3  w_e      = sum over the edge's 4 faces of +-(h/2)(y(c1) + y(c2))
#include "common.h"

namespace pyib {
namespace {

struct Lev {
  int nx, ny, nz;
  int64_t cell_off, e0, e1, e2;
  double h;
  double inv_nz, inv_ny, inv_nz1, inv_ny1;   // reciprocals for the index decode
};

struct Levels {
  int n;
  Lev l[PYIB_MAX_LEVELS];
};

Levels levels_of(const pyib_grid* g) {
  Levels L;
  L.n = g->nlevels;
  for (int i = 0; i < g->nlevels; ++i) {
    const pyib_level& s = g->lev[i];
    L.l[i] = Lev{s.nx, s.ny, s.nz, s.cell_off, s.e_off[0], s.e_off[1], s.e_off[2], s.h,
                 1.0 / s.nz, 1.0 / s.ny, 1.0 / (s.nz + 1), 1.0 / (s.ny + 1)};
  }
  return L;
}

__device__ __forceinline__ int level_of_cell(const Levels& L, int64_t c) {
  int k = 0;
  while (k + 1 < L.n && c >= L.l[k + 1].cell_off) ++k;
  return k;
}
__device__ __forceinline__ int level_of_edge(const Levels& L, int64_t e) {
  int k = 0;
  while (k + 1 < L.n && e >= L.l[k + 1].e0) ++k;
  return k;
}

// q / d for 0 <= q < 2^31 through a double reciprocal, corrected to the exact quotient
__device__ __forceinline__ int qdiv(int q, int d, double inv) {
  int t = (int)((double)q * inv);
  const int r = q - t * d;
  t += (r >= d) ? 1 : 0;
  t -= (r < 0) ? 1 : 0;
  return t;
}
// (i, j, k) of offset q in a C-ordered (., ny, nz) block
__device__ __forceinline__ void decode(int q, int ny, int nz, double inv_ny, double inv_nz, int& i, int& j, int& k) {
  const int t = qdiv(q, nz, inv_nz);
  k = q - t * nz;
  i = qdiv(t, ny, inv_ny);
  j = t - i * ny;
}

// ---------------------------------------------------------------- G1
template <typename T>
__global__ void k_sg1(int64_t ncell, Levels L, const uint8_t* __restrict__ reg, const T* __restrict__ p,
                      T* __restrict__ vc) {
  for (int64_t c = xcd_block() * blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    if (!reg[c]) continue;
    const Lev& v = L.l[level_of_cell(L, c)];
    int i, j, k;
    decode((int)(c - v.cell_off), v.ny, v.nz, v.inv_ny, v.inv_nz, i, j, k);
    const int ny1 = v.ny + 1, nz1 = v.nz + 1;
    // x-edges (nx, ny+1, nz+1), y-edges (nx+1, ny, nz+1), z-edges (nx+1, ny+1, nz)
    auto PX = [&](int a, int b, int d) { return p[v.e0 + ((int64_t)a * ny1 + b) * nz1 + d]; };
    auto PY = [&](int a, int b, int d) { return p[v.e1 + ((int64_t)a * v.ny + b) * nz1 + d]; };
    auto PZ = [&](int a, int b, int d) { return p[v.e2 + ((int64_t)a * ny1 + b) * v.nz + d]; };
    const T ux0 = (PY(i, j, k) - PY(i, j, k + 1)) + (PZ(i, j + 1, k) - PZ(i, j, k));
    const T ux1 = (PY(i + 1, j, k) - PY(i + 1, j, k + 1)) + (PZ(i + 1, j + 1, k) - PZ(i + 1, j, k));
    const T uy0 = (PX(i, j, k + 1) - PX(i, j, k)) - (PZ(i + 1, j, k) - PZ(i, j, k));
    const T uy1 = (PX(i, j + 1, k + 1) - PX(i, j + 1, k)) - (PZ(i + 1, j + 1, k) - PZ(i, j + 1, k));
    const T uz0 = (PY(i + 1, j, k) - PY(i, j, k)) - (PX(i, j + 1, k) - PX(i, j, k));
    const T uz1 = (PY(i + 1, j, k + 1) - PY(i, j, k + 1)) - (PX(i, j + 1, k + 1) - PX(i, j, k + 1));
    const T cf = (T)(1.0 / (2.0 * v.h * v.h));
    vc[c] = (ux0 + ux1) * cf;
    vc[ncell + c] = (uy0 + uy1) * cf;
    vc[2 * ncell + c] = (uz0 + uz1) * cf;
  }
}

// composite rows in sliced ELL, one wavefront per slice of 64 rows (lane = row):
// vc_d(row) = sum_j coef_d * p[col]
template <typename T>
__global__ void k_sell_g1(int64_t nslice, int64_t ncell, const int32_t* __restrict__ row,
                          const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                          const int32_t* __restrict__ col, const T* __restrict__ coef, const T* __restrict__ p,
                          T* __restrict__ vc) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t s = blockIdx.x * wpb + (threadIdx.x >> 6); s < nslice; s += (int64_t)gridDim.x * wpb) {
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 4
    for (int j = 0; j < W; ++j) {
      const int64_t e = b + (int64_t)j * 64;
      const T x = p[col[e]];
      a0 += coef[3 * e] * x; a1 += coef[3 * e + 1] * x; a2 += coef[3 * e + 2] * x;
    }
    const int r = row[s * 64 + lane];
    if (r >= 0) { vc[r] = a0; vc[ncell + r] = a1; vc[2 * ncell + r] = a2; }
  }
}

// ----------------------------------------------------------------- V
template <typename T>
__global__ void k_sv(int64_t ncell, Levels L, const uint8_t* __restrict__ reg, T rou, T visc, T hdt,
                     const T* __restrict__ x, T* __restrict__ y) {
  for (int64_t c = xcd_block() * blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    if (!reg[c]) continue;
    const Lev& v = L.l[level_of_cell(L, c)];
    const int64_t si = (int64_t)v.ny * v.nz, sj = v.nz;
    const T kA = (T)v.h * visc;
    const T hv = hdt * (T)(1.0 / (v.h * v.h * v.h));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T* xd = x + d * ncell;
      const T xc = xd[c];
      T a = kA * (xd[c + si] - xc) + kA * (xd[c + sj] - xc);
      a = a + kA * (xd[c + 1] - xc);
      T b = kA * (xc - xd[c - si]) + kA * (xc - xd[c - sj]);
      b = b + kA * (xc - xd[c - 1]);
      y[d * ncell + c] = rou * xc - hv * (a - b);
    }
  }
}

template <typename T>
__global__ void k_sell_v(int64_t nslice, int64_t ncell, const int32_t* __restrict__ row,
                         const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                         const int32_t* __restrict__ col, const T* __restrict__ coef, T rou, T kv,
                         const T* __restrict__ x, T* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  for (int64_t s = blockIdx.x * wpb + (threadIdx.x >> 6); s < nslice; s += (int64_t)gridDim.x * wpb) {
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 4
    for (int j = 0; j < W; ++j) {
      const int64_t e = b + (int64_t)j * 64;
      const int64_t c = col[e];
      const T cf = coef[e];
      a0 += cf * x[c]; a1 += cf * x[ncell + c]; a2 += cf * x[2 * ncell + c];
    }
    const int r = row[s * 64 + lane];
    if (r >= 0) {
      y[r] = rou * x[r] - kv * a0;
      y[ncell + r] = rou * x[ncell + r] - kv * a1;
      y[2 * ncell + r] = rou * x[2 * ncell + r] - kv * a2;
    }
  }
}

// ---------------------------------------------------------------- G3
// w_e of one regular edge e of level v
template <typename T>
__device__ __forceinline__ T sg3_value(const Lev& v, int64_t e, int64_t ncell, const T* __restrict__ y) {
  const int64_t si = (int64_t)v.ny * v.nz, sj = v.nz;
  const T hh = (T)(0.5 * v.h);
  const T* yx = y;
  const T* yy = y + ncell;
  const T* yz = y + 2 * ncell;
  int dir, i, j, k;
  if (e < v.e1) {
    dir = 0;
    decode((int)(e - v.e0), v.ny + 1, v.nz + 1, v.inv_ny1, v.inv_nz1, i, j, k);
  } else if (e < v.e2) {
    dir = 1;
    decode((int)(e - v.e1), v.ny, v.nz + 1, v.inv_ny, v.inv_nz1, i, j, k);
  } else {
    dir = 2;
    decode((int)(e - v.e2), v.ny + 1, v.nz, v.inv_ny1, v.inv_nz, i, j, k);
  }
  const int64_t c = v.cell_off + ((int64_t)i * v.ny + j) * v.nz + k;   // cell (i,j,k) of this level
  T val;
  if (dir == 0) {
    // -F_y(i,j,k) + F_y(i,j,k-1) + F_z(i,j,k) - F_z(i,j-1,k)
    const T fy0 = hh * (yy[c - sj] + yy[c]);
    const T fy1 = hh * (yy[c - sj - 1] + yy[c - 1]);
    const T fz0 = hh * (yz[c - 1] + yz[c]);
    const T fz1 = hh * (yz[c - sj - 1] + yz[c - sj]);
    val = (fy1 - fy0) + (fz0 - fz1);
  } else if (dir == 1) {
    // F_x(i,j,k) - F_x(i,j,k-1) - F_z(i,j,k) + F_z(i-1,j,k)
    const T fx0 = hh * (yx[c - si] + yx[c]);
    const T fx1 = hh * (yx[c - si - 1] + yx[c - 1]);
    const T fz0 = hh * (yz[c - 1] + yz[c]);
    const T fz1 = hh * (yz[c - si - 1] + yz[c - si]);
    val = (fx0 - fx1) + (fz1 - fz0);
  } else {
    // -F_x(i,j,k) + F_x(i,j-1,k) + F_y(i,j,k) - F_y(i-1,j,k)
    const T fx0 = hh * (yx[c - si] + yx[c]);
    const T fx1 = hh * (yx[c - si - sj] + yx[c - sj]);
    const T fy0 = hh * (yy[c - sj] + yy[c]);
    const T fy1 = hh * (yy[c - si - sj] + yy[c - si]);
    val = (fx1 - fx0) + (fy0 - fy1);
  }
  return val;
}

// block partials of p.w over the rows this kernel writes; each block takes chunks of
// kBlock*kPerThread consecutive edges (XCD-remapped), kPerThread coalesced rows per thread
template <typename T>
__global__ void k_sg3(int64_t nedge, int64_t ncell, Levels L, const uint8_t* __restrict__ reg,
                      const T* __restrict__ y, T* __restrict__ w, const T* __restrict__ p,
                      double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double dot = 0.0;
  const int64_t chunk = (int64_t)blockDim.x * kPerThread;
  for (int64_t base = xcd_block() * chunk; base < nedge; base += (int64_t)gridDim.x * chunk) {
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int64_t e = base + (int64_t)u * blockDim.x + threadIdx.x;
      if (e < nedge && reg[e]) {
        const T val = sg3_value(L.l[level_of_edge(L, e)], e, ncell, y);
        w[e] = val;
        if (p) dot += (double)p[e] * (double)val;
      }
    }
  }
  if (partials) {
    double r = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
  }
}

template <typename T>
__global__ void k_sell_g3(int64_t nslice, int64_t ncell, const int32_t* __restrict__ row,
                          const int32_t* __restrict__ width, const int64_t* __restrict__ sptr,
                          const int32_t* __restrict__ col, const T* __restrict__ coef,
                          const uint8_t* __restrict__ mask, const T* __restrict__ y, T* __restrict__ w,
                          const T* __restrict__ p, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  const int lane = threadIdx.x & 63;
  const int64_t wpb = blockDim.x >> 6;
  double dot = 0.0;
  for (int64_t s = blockIdx.x * wpb + (threadIdx.x >> 6); s < nslice; s += (int64_t)gridDim.x * wpb) {
    const int W = width[s];
    const int64_t b = sptr[s] + lane;
    T acc = 0;
#pragma unroll 4
    for (int j = 0; j < W; ++j) {
      const int64_t e = b + (int64_t)j * 64;
      const int64_t c = col[e];
      T t = coef[3 * e] * y[c] + coef[3 * e + 1] * y[ncell + c];
      t = t + coef[3 * e + 2] * y[2 * ncell + c];
      acc += t;
    }
    const int r = row[s * 64 + lane];
    if (r >= 0) {
      if (mask && mask[r]) acc = T(0);
      w[r] = acc;
      if (p) dot += (double)p[r] * (double)acc;
    }
  }
  if (partials) {
    double rr = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = rr;
  }
}

// sum of two partial sets (nb1 from the composite rows, nb2 from the stencil rows)
__global__ void k_alpha2(const double* __restrict__ pa, int na, const double* __restrict__ pb, int nb,
                         double* __restrict__ scal, int write_alpha, double* __restrict__ dot_out) {
  __shared__ double sh[kFinBlock / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < na; b += blockDim.x) acc += pa[b];
  double s1 = block_sum(acc, sh);
  acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += pb[b];
  double s2 = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double den = s1 + s2;
    if (dot_out) *dot_out = den;
    if (write_alpha) {
      scal[S_DEN] = den;
      scal[S_ALPHA] = scal[S_ETA] / (den != 0.0 ? den : __builtin_inf());
    }
  }
}

// ------------------------------------------------------ CG on the grid
template <typename T>
__global__ void k_gcg_update(int64_t n, const double* __restrict__ scal, T* __restrict__ x, T* __restrict__ r,
                             const T* __restrict__ p, const T* __restrict__ w, const T* __restrict__ prec,
                             double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  const T alf = (T)scal[S_ALPHA];
  double acc = 0.0;
  const int64_t chunk = (int64_t)blockDim.x * kPerThread;
  for (int64_t base = blockIdx.x * chunk; base < n; base += (int64_t)gridDim.x * chunk) {
    T xv[kPerThread], pv[kPerThread], rv[kPerThread], wv[kPerThread], mv[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
      if (i < n) { xv[u] = x[i]; pv[u] = p[i]; rv[u] = r[i]; wv[u] = w[i]; mv[u] = prec[i]; }
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int64_t i = base + (int64_t)u * blockDim.x + threadIdx.x;
      if (i < n) {
        x[i] = xv[u] + alf * pv[u];
        const T ri = rv[u] - alf * wv[u];
        r[i] = ri;
        acc += (double)(ri * mv[u]) * (double)ri;
      }
    }
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
__global__ void k_gcg_beta(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
  __shared__ double sh[kFinBlock / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[b];
  double en = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    scal[S_ETA_NEW] = en;
    const double eta = scal[S_ETA];
    scal[S_BETA] = en / (eta != 0.0 ? eta : __builtin_inf());
  }
}
template <typename T>
__global__ void k_gcg_dir(int64_t n, const double* __restrict__ scal, const T* __restrict__ r,
                          const T* __restrict__ prec, T* __restrict__ p) {
  const T beta = (T)scal[S_BETA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = r[i] * prec[i] + beta * p[i];
}
__global__ void k_gcg_shift(double* __restrict__ scal) { scal[S_ETA] = scal[S_ETA_NEW]; }

template <typename T>
__global__ void k_gcg_begin(int64_t n, const T* __restrict__ res, const T* __restrict__ s,
                            const uint8_t* __restrict__ mask, const T* __restrict__ prec, T* __restrict__ r,
                            T* __restrict__ x, T* __restrict__ p, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double rr = 0, pr = 0, xx = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = mask[i] ? T(0) : res[i];
    const T xi = s[i];
    const T pi = ri * prec[i];
    r[i] = ri; x[i] = xi; p[i] = pi;
    rr += (double)ri * (double)ri; pr += (double)pi * (double)ri; xx += (double)xi * (double)xi;
  }
  double a = block_sum(rr, sh), b = block_sum(pr, sh), c = block_sum(xx, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a; partials[kRedGrid + blockIdx.x] = b; partials[2 * kRedGrid + blockIdx.x] = c;
  }
}
__global__ void k_gcg_begin_fin(const double* __restrict__ partials, int nb, double tol1, double tol2,
                                double* __restrict__ scal) {
  __shared__ double sh[kFinBlock / 64];
  double v[3];
  for (int k = 0; k < 3; ++k) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[k * kRedGrid + b];
    v[k] = block_sum(acc, sh);
  }
  if (threadIdx.x == 0) {
    const double rr = v[0], eta = v[1], xnorm = v[2] + 1e-16;
    scal[S_L2I] = rr; scal[S_ETA] = eta;
    scal[S_ERR] = xnorm * (tol1 * tol1) + (tol2 * tol2) * eta;
  }
}
template <typename T>
__global__ void k_gsub(int64_t n, const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}
template <typename T>
__global__ void k_gtrue_res(int64_t n, const T* __restrict__ res, const T* __restrict__ img,
                            const uint8_t* __restrict__ mask, const T* __restrict__ prec, T* __restrict__ tr,
                            double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double a = 0, b = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T t = mask[i] ? T(0) : res[i] - img[i];
    tr[i] = t;
    a += (double)(t * prec[i]) * (double)t;
    b += (double)t * (double)t;
  }
  double ra = block_sum(a, sh), rb = block_sum(b, sh);
  if (threadIdx.x == 0) { partials[blockIdx.x] = ra; partials[kRedGrid + blockIdx.x] = rb; }
}

// signed gather between numberings: out[i] = idx[i] >= 0 ? sigma[i] * in[idx[i]] : 0
template <typename A, typename B>
__global__ void k_permute(int64_t n, const int64_t* __restrict__ idx, const int8_t* __restrict__ sigma,
                          const A* __restrict__ in, B* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx[i];
    out[i] = s >= 0 ? (B)((A)sigma[i] * in[s]) : B(0);
  }
}

// ------------------------------------------------------- launchers
inline int slice_grid(int64_t nslice) { return grid_for(nslice * 64); }
inline int slice_red_grid(int64_t nslice) { return red_grid_for(nslice * 64); }

// w = A p on the grid; the p.w partials go to ws->partials [0, nb1) and [kRedGrid, kRedGrid+nb2)
template <typename T>
int gflow(const pyib_grid* g, const pyib_work* ws, double rou, double visc, double hdt, const T* p, T* w,
          bool want_dot, int* nb1, int* nb2, hipStream_t st) {
  const Levels L = levels_of(g);
  T* ca = static_cast<T*>(ws->cell_a);
  T* cb = static_cast<T*>(ws->cell_b);
  k_sg1<T><<<grid_for(g->ncell), kBlock, 0, st>>>(g->ncell, L, g->cell_reg, p, ca);
  if (g->g1_nslice > 0)
    k_sell_g1<T><<<slice_grid(g->g1_nslice), kBlock, 0, st>>>(g->g1_nslice, g->ncell, g->g1_row, g->g1_width,
                                                             g->g1_sptr, g->g1_col, (const T*)g->g1_coef, p, ca);
  k_sv<T><<<grid_for(g->ncell), kBlock, 0, st>>>(g->ncell, L, g->cell_reg, (T)rou, (T)visc, (T)hdt, ca, cb);
  if (g->v_nslice > 0)
    k_sell_v<T><<<slice_grid(g->v_nslice), kBlock, 0, st>>>(g->v_nslice, g->ncell, g->v_row, g->v_width, g->v_sptr,
                                                           g->v_col, (const T*)g->v_coef, (T)rou, (T)(hdt * visc),
                                                           ca, cb);
  const int n1 = slice_red_grid(g->g3_nslice), n2 = red_grid_for(g->nedge);
  if (g->g3_nslice > 0)
    k_sell_g3<T><<<n1, kBlock, 0, st>>>(g->g3_nslice, g->ncell, g->g3_row, g->g3_width, g->g3_sptr, g->g3_col,
                                        (const T*)g->g3_coef, g->dir, cb, w, want_dot ? p : nullptr,
                                        want_dot ? ws->partials : nullptr);
  k_sg3<T><<<n2, kBlock, 0, st>>>(g->nedge, g->ncell, L, g->edge_reg, cb, w, want_dot ? p : nullptr,
                                  want_dot ? ws->partials + kRedGrid : nullptr);
  *nb1 = g->g3_nslice > 0 ? n1 : 0;
  *nb2 = n2;
  return 0;
}

template <typename T>
int gcg_iterate(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, double rou, double visc, double hdt,
                int n_iter, int skip_last, hipStream_t st) {
  T* x = static_cast<T*>(cg->x);
  T* r = static_cast<T*>(cg->r);
  T* p = static_cast<T*>(cg->p);
  T* w = static_cast<T*>(cg->w);
  const T* prec = static_cast<const T*>(cg->prec);
  const int64_t n = g->nedge;
  const int nb = red_grid_for(n);
  // the update's partials use a third row so they never alias the dot of A p
  double* upart = ws->partials + 2 * kRedGrid;
  for (int it = 0; it < n_iter; ++it) {
    int nb1, nb2;
    gflow<T>(g, ws, rou, visc, hdt, p, w, true, &nb1, &nb2, st);
    k_alpha2<<<1, kFinBlock, 0, st>>>(ws->partials, nb1, ws->partials + kRedGrid, nb2, cg->scal, 1, nullptr);
    k_gcg_update<T><<<nb, kBlock, 0, st>>>(n, cg->scal, x, r, p, w, prec, upart);
    k_gcg_beta<<<1, kFinBlock, 0, st>>>(upart, nb, cg->scal);
    if (!(skip_last && it == n_iter - 1)) {
      k_gcg_dir<T><<<grid_for(n), kBlock, 0, st>>>(n, cg->scal, r, prec, p);
      k_gcg_shift<<<1, 1, 0, st>>>(cg->scal);
    }
  }
  return check_launch("pyib_grid_cg_iterate");
}

}  // namespace
}  // namespace pyib

extern "C" {
using namespace pyib;

int pyib_grid_flow_mat(const pyib_grid* g, const pyib_work* ws, double rou, double visc, double hdt, const void* p,
                       void* w, double* dot_out, void* stream) {
  auto st = S(stream);
  int nb1 = 0, nb2 = 0;
  if (g->dtype == PYIB_F64)
    gflow<double>(g, ws, rou, visc, hdt, (const double*)p, (double*)w, dot_out != nullptr, &nb1, &nb2, st);
  else if (g->dtype == PYIB_F32)
    gflow<float>(g, ws, rou, visc, hdt, (const float*)p, (float*)w, dot_out != nullptr, &nb1, &nb2, st);
  else
    return fail_msg("pyib_grid_flow_mat: unknown dtype");
  if (dot_out) k_alpha2<<<1, kFinBlock, 0, st>>>(ws->partials, nb1, ws->partials + kRedGrid, nb2, nullptr, 0, dot_out);
  return check_launch("pyib_grid_flow_mat");
}

int pyib_grid_cg_begin(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, const void* res, const void* s,
                       double tol1, double tol2, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(g->nedge);
  if (g->dtype == PYIB_F64)
    k_gcg_begin<double><<<nb, kBlock, 0, st>>>(g->nedge, (const double*)res, (const double*)s, g->dir,
        (const double*)cg->prec, (double*)cg->r, (double*)cg->x, (double*)cg->p, ws->partials);
  else
    k_gcg_begin<float><<<nb, kBlock, 0, st>>>(g->nedge, (const float*)res, (const float*)s, g->dir,
        (const float*)cg->prec, (float*)cg->r, (float*)cg->x, (float*)cg->p, ws->partials);
  k_gcg_begin_fin<<<1, kFinBlock, 0, st>>>(ws->partials, nb, tol1, tol2, cg->scal);
  return check_launch("pyib_grid_cg_begin");
}

int pyib_grid_cg_iterate(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, double rou, double visc,
                         double hdt, int32_t n_iter, int32_t skip_last_direction, void* stream) {
  if (n_iter < 0) return fail_msg("pyib_grid_cg_iterate: negative n_iter");
  if (g->dtype == PYIB_F64)
    return gcg_iterate<double>(g, ws, cg, rou, visc, hdt, n_iter, skip_last_direction, S(stream));
  if (g->dtype == PYIB_F32)
    return gcg_iterate<float>(g, ws, cg, rou, visc, hdt, n_iter, skip_last_direction, S(stream));
  return fail_msg("pyib_grid_cg_iterate: unknown dtype");
}

int pyib_grid_cg_direction(const pyib_grid* g, const pyib_cg* cg, void* stream) {
  auto st = S(stream);
  if (g->dtype == PYIB_F64)
    k_gcg_dir<double><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, cg->scal, (const double*)cg->r,
                                                             (const double*)cg->prec, (double*)cg->p);
  else
    k_gcg_dir<float><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, cg->scal, (const float*)cg->r,
                                                            (const float*)cg->prec, (float*)cg->p);
  k_gcg_shift<<<1, 1, 0, st>>>(cg->scal);
  return check_launch("pyib_grid_cg_direction");
}

int pyib_grid_cg_true_residual(const pyib_grid* g, const pyib_work* ws, const pyib_cg* cg, const void* res,
                               const void* s, void* tmp, void* tr, double rou, double visc, double hdt, double* out2,
                               void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(g->nedge);
  int nb1, nb2;
  if (g->dtype == PYIB_F64) {
    k_gsub<double><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, (const double*)cg->x, (const double*)s,
                                                          (double*)tmp);
    gflow<double>(g, ws, rou, visc, hdt, (const double*)tmp, (double*)tr, false, &nb1, &nb2, st);
    k_gtrue_res<double><<<nb, kBlock, 0, st>>>(g->nedge, (const double*)res, (const double*)tr, g->dir,
                                               (const double*)cg->prec, (double*)tr, ws->partials);
  } else {
    k_gsub<float><<<grid_for(g->nedge), kBlock, 0, st>>>(g->nedge, (const float*)cg->x, (const float*)s, (float*)tmp);
    gflow<float>(g, ws, rou, visc, hdt, (const float*)tmp, (float*)tr, false, &nb1, &nb2, st);
    k_gtrue_res<float><<<nb, kBlock, 0, st>>>(g->nedge, (const float*)res, (const float*)tr, g->dir,
                                              (const float*)cg->prec, (float*)tr, ws->partials);
  }
  k_finalize<<<1, kFinBlock, 0, st>>>(ws->partials, nb, 2, out2);
  return check_launch("pyib_grid_cg_true_residual");
}

int pyib_permute(int dtype_in, const void* in, int dtype_out, void* out, int64_t n_out, const int64_t* idx,
                 const int8_t* sigma, void* stream) {
  auto st = S(stream);
  const int gr = grid_for(n_out);
  if (dtype_in == PYIB_F64 && dtype_out == PYIB_F64)
    k_permute<double, double><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const double*)in, (double*)out);
  else if (dtype_in == PYIB_F64 && dtype_out == PYIB_F32)
    k_permute<double, float><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const double*)in, (float*)out);
  else if (dtype_in == PYIB_F32 && dtype_out == PYIB_F64)
    k_permute<float, double><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const float*)in, (double*)out);
  else
    k_permute<float, float><<<gr, kBlock, 0, st>>>(n_out, idx, sigma, (const float*)in, (float*)out);
  return check_launch("pyib_permute");
}

}  // extern "C"
