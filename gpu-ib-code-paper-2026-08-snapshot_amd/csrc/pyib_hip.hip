// pyib_hip.hip -- gfx950 kernels and C ABI of PyIB's time-step hot path.
//
// Operators work on PyIB's global (unstructured) numbering through incidence
// tables built once on the host: every output element gathers its inputs, so
// there are no atomics anywhere and every result is bitwise reproducible.
// Reductions are two-stage (fixed block partition -> one finalising block),
// deterministic for a given size, and always accumulate in double.
//
// Summation orders follow the reference CPU backend where that is free: each
// gather visits its links in link order and keeps the reference's split into
// first-owner and second-owner sums (NumPy bincount(c1) - bincount(c2)).
#include <hipcub/hipcub.hpp>
#include <cstring>
#include <limits>
#include <mutex>

#include "common.h"

namespace pyib {

thread_local char g_err[512] = "";

int fail(const char* where, hipError_t e) {
  std::snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
  return 1;
}
int fail_msg(const char* msg) {
  std::snprintf(g_err, sizeof(g_err), "%s", msg);
  return 2;
}

// Sum nsum rows of `nb` partials each: out[k] = sum_b partials[k*kRedGrid + b].
__global__ void k_finalize(const double* __restrict__ partials, int nb, int nsum, double* __restrict__ out) {
  __shared__ double sh[kFinBlock / 64];
  for (int k = 0; k < nsum; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partials[k * kRedGrid + b];
    double r = block_sum(v, sh);
    if (threadIdx.x == 0) out[k] = r;
  }
}
__global__ void k_finalize_max(const double* __restrict__ partials, int nb, int nsum, double* __restrict__ out) {
  __shared__ double sh[kFinBlock / 64];
  for (int k = 0; k < nsum; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v = nanmax(v, partials[k * kRedGrid + b]);
    double r = block_max(v, sh);
    if (threadIdx.x == 0) out[k] = r;
  }
}

// ---------------------------------------------------------------- operators
// curl: U[f] = sum over the face's links of sign * s[e]   (ib_core_np.py:133-134)
template <typename T>
__global__ void k_curl(int64_t nface, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                       const T* __restrict__ s, T* __restrict__ U) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nface; f += (int64_t)gridDim.x * blockDim.x) {
    T acc = T(0);
    for (int j = ptr[f]; j < ptr[f + 1]; ++j) acc += sload(s, idx[j]);
    U[f] = acc;
  }
}

// rot: w[e] = sum over the edge's links of sign * F[f]   (ib_core_np.py:129-131)
// optional: zero on dir_mask, and block partials of p.w
template <typename T>
__global__ void k_rot(int64_t nedge, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                      const T* __restrict__ F, T* __restrict__ w, const uint8_t* __restrict__ mask,
                      const T* __restrict__ p, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double dot = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nedge; e += (int64_t)gridDim.x * blockDim.x) {
    T acc = T(0);
    for (int j = ptr[e]; j < ptr[e + 1]; ++j) acc += sload(F, idx[j]);
    if (mask && mask[e]) acc = T(0);
    w[e] = acc;
    if (p) dot += (double)p[e] * (double)acc;
  }
  if (partials) {
    double r = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
  }
}

// interp_f2c: vc[c,d] = (sum_{c=c1} U rf1[:,d] - sum_{c=c2} U rf2[:,d]) * vol_ci  (ib_core_np.py:150-157)
template <typename T>
__global__ void k_interp(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                         const int32_t* __restrict__ face, const T* __restrict__ rf1, const T* __restrict__ rf2,
                         const T* __restrict__ vol_ci, const T* __restrict__ U, T* __restrict__ vc) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    T a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
    const int j0 = ptr[c], jm = mid[c], j1 = ptr[c + 1];
    for (int j = j0; j < jm; ++j) {
      const int f = face[j];
      const T u = U[f];
      a0 += u * rf1[3 * f]; a1 += u * rf1[3 * f + 1]; a2 += u * rf1[3 * f + 2];
    }
    for (int j = jm; j < j1; ++j) {
      const int f = face[j];
      const T u = U[f];
      b0 += u * rf2[3 * f]; b1 += u * rf2[3 * f + 1]; b2 += u * rf2[3 * f + 2];
    }
    const T v = vol_ci[c];
    vc[3 * c] = (a0 - b0) * v; vc[3 * c + 1] = (a1 - b1) * v; vc[3 * c + 2] = (a2 - b2) * v;
  }
}

// viscous block of flow_mat (ib_core_np.py:171-174), three components:
// y = rou x - (hdt vol_ci) cdiv((AoL visc) cgrad0 x), ghost values zero.
template <typename T>
__global__ void k_visc(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                       const int32_t* __restrict__ face, const int32_t* __restrict__ other,
                       const T* __restrict__ aol, const T* __restrict__ vol_ci, T visc, T rou, T hdt,
                       const T* __restrict__ x, T* __restrict__ y) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    const T x0 = x[3 * c], x1 = x[3 * c + 1], x2 = x[3 * c + 2];
    T a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
    const int j0 = ptr[c], jm = mid[c], j1 = ptr[c + 1];
    for (int j = j0; j < jm; ++j) {
      const T k = aol[face[j]] * visc;
      const int o = other[j];
      T o0 = 0, o1 = 0, o2 = 0;
      if (o < ncell) { o0 = x[3 * o]; o1 = x[3 * o + 1]; o2 = x[3 * o + 2]; }
      a0 += k * (o0 - x0); a1 += k * (o1 - x1); a2 += k * (o2 - x2);
    }
    for (int j = jm; j < j1; ++j) {
      const T k = aol[face[j]] * visc;
      const int o = other[j];
      b0 += k * (x0 - x[3 * o]); b1 += k * (x1 - x[3 * o + 1]); b2 += k * (x2 - x[3 * o + 2]);
    }
    const T hv = hdt * vol_ci[c];
    y[3 * c] = rou * x0 - hv * (a0 - b0);
    y[3 * c + 1] = rou * x1 - hv * (a1 - b1);
    y[3 * c + 2] = rou * x2 - hv * (a2 - b2);
  }
}

// integr_c2f: F[f] = vc[c1].rf1 - vc[c2].rf2   (ib_core_np.py:144-148)
template <typename T>
__global__ void k_integr(int64_t nface, const int32_t* __restrict__ f2c, const T* __restrict__ rf1,
                         const T* __restrict__ rf2, const T* __restrict__ vc, T* __restrict__ F) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nface; f += (int64_t)gridDim.x * blockDim.x) {
    const int c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
    T a = vc[3 * c1] * rf1[3 * f] + vc[3 * c1 + 1] * rf1[3 * f + 1];
    a = a + vc[3 * c1 + 2] * rf1[3 * f + 2];
    if (c2 >= 0) {
      T b = vc[3 * c2] * rf2[3 * f] + vc[3 * c2 + 1] * rf2[3 * f + 1];
      b = b + vc[3 * c2 + 2] * rf2[3 * f + 2];
      a = a - b;
    }
    F[f] = a;
  }
}

// cdiv: out[c] = sum_{c=c1} F - sum_{c=c2} F   (ib_core_np.py:136-139)
template <typename T>
__global__ void k_cdiv(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                       const int32_t* __restrict__ face, const T* __restrict__ F, const T* __restrict__ scale,
                       T* __restrict__ out) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    T a = 0, b = 0;
    for (int j = ptr[c]; j < mid[c]; ++j) a += F[face[j]];
    for (int j = mid[c]; j < ptr[c + 1]; ++j) b += F[face[j]];
    out[c] = scale ? scale[c] * (a - b) : (a - b);
  }
}

// ------------------------------------------------------------ time step
template <typename T>
__global__ void k_copy_inlet(int64_t n, const T* __restrict__ s_in, T* __restrict__ s_out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s_out[i] = s_in[i];
}
template <typename T>
__global__ void k_set_inlet(int64_t n, const int32_t* __restrict__ e, const T* __restrict__ v, T* __restrict__ s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s[e[i]] = v[i];
}

// ghost cells of calc_vel (ib_driver_np.py:275-291)
template <typename T>
__global__ void k_ghost(int64_t nghost, int64_t ncell, const int32_t* __restrict__ gface,
                        const int32_t* __restrict__ gcell, const int32_t* __restrict__ gkind,
                        const T* __restrict__ gnf, const T* __restrict__ gaf2, const T* __restrict__ U,
                        T lid0, T lid1, T lid2, T* __restrict__ vc) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nghost; k += (int64_t)gridDim.x * blockDim.x) {
    const int c = gcell[k];
    const T v0 = vc[3 * c], v1 = vc[3 * c + 1], v2 = vc[3 * c + 2];
    const T n0 = gnf[3 * k], n1 = gnf[3 * k + 1], n2 = gnf[3 * k + 2];
    T g0, g1, g2;
    switch (gkind[k]) {
      case 0: {  // inlet: face velocity normal flux only
        const T u = U[gface[k]], a2 = gaf2[k];
        g0 = (u * n0) / a2; g1 = (u * n1) / a2; g2 = (u * n2) / a2;
        break;
      }
      case 1: {  // zero gradient with the normal flux imposed
        const T u = U[gface[k]], a2 = gaf2[k];
        T d = v0 * n0 + v1 * n1;
        d = d + v2 * n2;
        const T q = u - d;
        g0 = v0 + (q * n0) / a2; g1 = v1 + (q * n1) / a2; g2 = v2 + (q * n2) / a2;
        break;
      }
      case 2: g0 = -v0; g1 = -v1; g2 = -v2; break;                                   // no-slip wall
      case 3: g0 = T(2) * lid0 - v0; g1 = T(2) * lid1 - v1; g2 = T(2) * lid2 - v2; break;  // moving lid
      default: g0 = v0; g1 = v1; g2 = v2; break;                                      // spanwise copy
    }
    const int64_t g = ncell + k;
    vc[3 * g] = g0; vc[3 * g + 1] = g1; vc[3 * g + 2] = g2;
  }
}

// conv_diff (ib_driver_np.py:293-303) fused with the momentum residual of step()
// (ib_driver_np.py:458, 466, 470): out = first - k_conv*conv + k_diff*diff (+ k_force*F)
template <typename T>
__global__ void k_momentum(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                           const int32_t* __restrict__ face, const int32_t* __restrict__ other,
                           const T* __restrict__ aol, const T* __restrict__ vol_ci, const T* __restrict__ U,
                           const T* __restrict__ vc, T visc, const T* __restrict__ base, T a_vc, T k_conv,
                           T k_diff, const T* __restrict__ force, T k_force, T* __restrict__ out) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    const T x0 = vc[3 * c], x1 = vc[3 * c + 1], x2 = vc[3 * c + 2];
    T ca0 = 0, ca1 = 0, ca2 = 0, cb0 = 0, cb1 = 0, cb2 = 0;
    T da0 = 0, da1 = 0, da2 = 0, db0 = 0, db1 = 0, db2 = 0;
    const int j0 = ptr[c], jm = mid[c], j1 = ptr[c + 1];
    for (int j = j0; j < jm; ++j) {  // c is the first owner: rd = v[o] - v[c]
      const int f = face[j], o = other[j];
      const T wf = T(0.5) * U[f];
      const T k = aol[f] * visc;
      const T o0 = vc[3 * o], o1 = vc[3 * o + 1], o2 = vc[3 * o + 2];
      ca0 += x0 * wf + o0 * wf; ca1 += x1 * wf + o1 * wf; ca2 += x2 * wf + o2 * wf;
      da0 += k * (o0 - x0); da1 += k * (o1 - x1); da2 += k * (o2 - x2);
    }
    for (int j = jm; j < j1; ++j) {  // c is the second owner: rd = v[c] - v[o]
      const int f = face[j], o = other[j];
      const T wf = T(0.5) * U[f];
      const T k = aol[f] * visc;
      const T o0 = vc[3 * o], o1 = vc[3 * o + 1], o2 = vc[3 * o + 2];
      cb0 += o0 * wf + x0 * wf; cb1 += o1 * wf + x1 * wf; cb2 += o2 * wf + x2 * wf;
      db0 += k * (x0 - o0); db1 += k * (x1 - o1); db2 += k * (x2 - o2);
    }
    const T v = vol_ci[c];
    const T conv[3] = {v * (ca0 - cb0), v * (ca1 - cb1), v * (ca2 - cb2)};
    const T diff[3] = {v * (da0 - db0), v * (da1 - db1), v * (da2 - db2)};
    const T xs[3] = {x0, x1, x2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T t = base ? base[3 * c + d] - a_vc * xs[d] : a_vc * xs[d];
      t = t - k_conv * conv[d];
      t = t + k_diff * diff[d];
      if (force) t = t + k_force * force[3 * c + d];
      out[3 * c + d] = t;
    }
  }
}

// ------------------------------------------------------------------ CG
template <typename T>
__global__ void k_cg_begin(int64_t n, const T* __restrict__ res_e, const T* __restrict__ s,
                           const uint8_t* __restrict__ mask, const T* __restrict__ prec, T* __restrict__ r,
                           T* __restrict__ x, T* __restrict__ p, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double rr = 0, pr = 0, xx = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = mask[i] ? T(0) : res_e[i];
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
__global__ void k_cg_begin_fin(const double* __restrict__ partials, int nb, double tol1, double tol2,
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
// alpha = eta / den (den != 0 else inf)     (ib_driver_np.py:389-392)
__global__ void k_cg_alpha(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
  __shared__ double sh[kFinBlock / 64];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[b];
  double den = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    scal[S_DEN] = den;
    scal[S_ALPHA] = scal[S_ETA] / (den != 0.0 ? den : __builtin_inf());
  }
}
// x += alpha p ; r -= alpha w ; eta_new partials = (r prec).r    (ib_driver_np.py:393-396)
template <typename T>
__global__ void k_cg_update(int64_t n, const double* __restrict__ scal, T* __restrict__ x, T* __restrict__ r,
                            const T* __restrict__ p, const T* __restrict__ w, const T* __restrict__ prec,
                            double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  const T alf = (T)scal[S_ALPHA];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = x[i] + alf * p[i];
    const T ri = r[i] - alf * w[i];
    r[i] = ri;
    const T z = ri * prec[i];
    acc += (double)z * (double)ri;
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
// eta_new and beta = eta_new / eta (eta != 0 else inf)   (ib_driver_np.py:396, 401)
__global__ void k_cg_beta(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
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
// p = r prec + beta p    (ib_driver_np.py:402)
template <typename T>
__global__ void k_cg_dir(int64_t n, const double* __restrict__ scal, const T* __restrict__ r,
                         const T* __restrict__ prec, T* __restrict__ p) {
  const T beta = (T)scal[S_BETA];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = r[i] * prec[i] + beta * p[i];
}
__global__ void k_cg_shift(double* __restrict__ scal) { scal[S_ETA] = scal[S_ETA_NEW]; }

// warm-start guess s + (s - s_old) * dt / dt_old   (ib_driver_np.py:461)
__global__ void k_extrapolate(int64_t n, const double* __restrict__ s, const double* __restrict__ s_old, double dt,
                              double dt_old, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = s[i] + ((s[i] - s_old[i]) * dt) / dt_old;
}

template <typename T>
__global__ void k_sub(int64_t n, const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}
// tr = res_e - img (0 on dir); partials of (tr prec).tr and tr.tr   (ib_driver_np.py:340, 406-409)
template <typename T>
__global__ void k_true_res(int64_t n, const T* __restrict__ res_e, const T* __restrict__ img,
                           const uint8_t* __restrict__ mask, const T* __restrict__ prec, T* __restrict__ tr,
                           double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double a = 0, b = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T t = mask[i] ? T(0) : res_e[i] - img[i];
    tr[i] = t;
    a += (double)(t * prec[i]) * (double)t;
    b += (double)t * (double)t;
  }
  double ra = block_sum(a, sh), rb = block_sum(b, sh);
  if (threadIdx.x == 0) { partials[blockIdx.x] = ra; partials[kRedGrid + blockIdx.x] = rb; }
}

// ---------------------------------------------------------- vector helpers
template <typename T>
__global__ void k_dot(int64_t n, const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
                      double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double t = (double)a[i] * (double)b[i];
    if (c) t *= (double)c[i];
    acc += t;
  }
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = r;
}
template <typename T>
__global__ void k_axpby(int64_t n, T a, const T* __restrict__ x, T b, const T* __restrict__ y, T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (y ? a * x[i] + b * y[i] : a * x[i]);
}
template <typename A, typename B>
__global__ void k_convert(int64_t n, const A* __restrict__ in, B* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (B)in[i];
}
template <typename A>
__global__ void k_add_into(int64_t n, const A* __restrict__ a, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + (double)a[i];
}

// ------------------------------------------------------ immersed boundary
// Roma-Peskin 3-point regularised delta   (ib_immersed.py:12-19)
__device__ __forceinline__ double delta_h3(double r, double h) {
  const double a = fabs(r / h);
  if (a <= 0.5) return (1.0 + sqrt(fmax(1.0 - 3.0 * a * a, 0.0))) / (3.0 * h);
  if (a <= 1.5) {
    const double q = 1.0 - a;
    return (5.0 - 3.0 * a - sqrt(fmax(1.0 - 3.0 * (q * q), 0.0))) / (6.0 * h);
  }
  return 0.0;
}

// one thread per (marker, tap); taps ordered like the reference's 5^3 window (ib_immersed.py:63-81)
__global__ void k_ib_stencil(int64_t nmark, const double* __restrict__ mk, double pr0, double pr1, double pr2,
                             double dx, double hh, int nx, int ny, int nz, int32_t* __restrict__ cell,
                             double* __restrict__ w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nmark * 27; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = i / 27;
    const int t = (int)(i % 27);
    const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
    const double mx = mk[3 * m], my = mk[3 * m + 1], mz = mk[3 * m + 2];
    const long ii = (long)rint((mx - pr0) / dx) + di;
    const long jj = (long)rint((my - pr1) / dx) + dj;
    const long kk = (long)rint((mz - pr2) / dx) + dk;
    const bool inb = ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz;
    if (!inb) { cell[i] = -1; w[i] = 0.0; continue; }
    const double gx = (pr0 + (double)ii * dx) - mx;
    const double gy = (pr1 + (double)jj * dx) - my;
    const double gz = (pr2 + (double)kk * dx) - mz;
    cell[i] = (int32_t)((ii * ny + jj) * nz + kk);
    w[i] = (delta_h3(gx, hh) * delta_h3(gy, hh)) * delta_h3(gz, hh);
  }
}

// vel[m] = sum_t vc[cell] (w hh^3)   (ib_immersed.py:87-93)
template <typename T>
__global__ void k_ib_interp(int64_t nmark, const int32_t* __restrict__ cell, const double* __restrict__ w,
                            double hh3, const T* __restrict__ vc, double* __restrict__ vel) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < nmark; m += (int64_t)gridDim.x * blockDim.x) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int t = 0; t < 27; ++t) {
      const int c = cell[27 * m + t];
      if (c < 0) continue;
      const double wt = w[27 * m + t] * hh3;
      a0 += (double)vc[3 * c] * wt; a1 += (double)vc[3 * c + 1] * wt; a2 += (double)vc[3 * c + 2] * wt;
    }
    vel[3 * m] = a0; vel[3 * m + 1] = a1; vel[3 * m + 2] = a2;
  }
}

constexpr int32_t kNoCell = 0x7fffffff;   // sort key of a tap outside the mesh (sorts last)

__global__ void k_ib_keys(int64_t n, const int32_t* __restrict__ cell, int32_t* __restrict__ keys,
                          int32_t* __restrict__ vals) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = cell[i];
    keys[i] = c < 0 ? kNoCell : c;
    vals[i] = (int32_t)i;
  }
}

// Segmented spreading over cell-sorted taps, reduced in LDS (ib_immersed.py:103-111).
// A block takes kBlock consecutive sorted taps: every lane gathers its tap's
// contribution flag[m] * (w * ds3[m]) into LDS, then each segment head (first tap
// of a cell's run) sums its run from LDS in marker order, the order NumPy's
// bincount adds in, and writes or accumulates the cell once.  A run that goes on
// past the block's taps is finished from global memory by the same head, so no
// cell is written by two blocks and no atomics are used: repeated runs are
// bitwise identical.

__device__ __forceinline__ void tap_contribution(int32_t v, const double* __restrict__ w,
                                                 const double* __restrict__ flag, const double* __restrict__ ds3,
                                                 double& c0, double& c1, double& c2) {
  const int m = v / 27;
  const double wt = w[v] * ds3[m];
  c0 = flag[3 * m] * wt; c1 = flag[3 * m + 1] * wt; c2 = flag[3 * m + 2] * wt;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_ib_spread(int64_t n, const int32_t* __restrict__ keys,
                                                      const int32_t* __restrict__ vals, const double* __restrict__ w,
                                                      const double* __restrict__ flag, const double* __restrict__ ds3,
                                                      double scale, int accumulate, T* __restrict__ out) {
  __shared__ int32_t skey[kBlock];
  __shared__ double sc[3][kBlock];
  const int64_t nchunk = (n + kBlock - 1) / kBlock;
  for (int64_t chunk = blockIdx.x; chunk < nchunk; chunk += gridDim.x) {   // block-uniform trip count
    const int64_t b0 = chunk * kBlock;
    const int t = threadIdx.x;
    const int64_t i = b0 + t;
    int32_t c = kNoCell;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    if (i < n) {
      c = keys[i];
      if (c != kNoCell) tap_contribution(vals[i], w, flag, ds3, c0, c1, c2);
    }
    skey[t] = c;
    sc[0][t] = c0; sc[1][t] = c1; sc[2][t] = c2;
    __syncthreads();
    const bool head = c != kNoCell && (t > 0 ? skey[t - 1] != c : (i == 0 || keys[i - 1] != c));
    if (head) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      int j = t;
      for (; j < kBlock && skey[j] == c; ++j) { s0 += sc[0][j]; s1 += sc[1][j]; s2 += sc[2][j]; }
      if (j == kBlock) {                      // the run continues into the next block's taps
        for (int64_t g = b0 + kBlock; g < n && keys[g] == c; ++g) {
          double d0, d1, d2;
          tap_contribution(vals[g], w, flag, ds3, d0, d1, d2);
          s0 += d0; s1 += d1; s2 += d2;
        }
      }
      if (accumulate) {
        out[3 * c] = (T)((double)out[3 * c] + scale * s0);
        out[3 * c + 1] = (T)((double)out[3 * c + 1] + scale * s1);
        out[3 * c + 2] = (T)((double)out[3 * c + 2] + scale * s2);
      } else {
        out[3 * c] = (T)(scale * s0); out[3 * c + 1] = (T)(scale * s1); out[3 * c + 2] = (T)(scale * s2);
      }
    }
    __syncthreads();                          // LDS is rewritten by the next chunk
  }
}

__global__ void k_ib_force(int64_t n, const double* __restrict__ vd, const double* __restrict__ vib, double dt,
                           double* __restrict__ f) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    f[i] = (vd[i] - vib[i]) / dt;
}

// ---------------------------------------------------------------- loads
// interior momenta over candidate cells (ib_interior.py:57-70)
template <typename T>
__global__ void k_interior(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ pos,
                           const double* __restrict__ vol, const T* __restrict__ vc, double cx, double cy,
                           double cz, double radius, double width, int fixed, const double* __restrict__ fwv,
                           double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    double wv;
    if (fixed) {
      wv = fwv[i];
    } else {
      const double dx = px - cx, dy = py - cy, dz = pz - cz;
      double ss = dx * dx + dy * dy;
      ss = ss + dz * dz;
      const double d = sqrt(ss);
      const double ind = fmin(fmax(((radius + 0.5 * width) - d) / width, 0.0), 1.0);
      wv = ind * vol[i];
    }
    const int c = idx[i];
    const double vx = (double)vc[3 * c], vy = (double)vc[3 * c + 1], vz = (double)vc[3 * c + 2];
    acc[0] += vx * wv; acc[1] += vy * wv; acc[2] += vz * wv;
    acc[3] += (py * vz - pz * vy) * wv; acc[4] += (pz * vx - px * vz) * wv; acc[5] += (px * vy - py * vx) * wv;
    acc[6] += wv;
  }
  for (int k = 0; k < 7; ++k) {
    double r = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partials[k * kRedGrid + blockIdx.x] = r;
  }
}
template <typename T>
__global__ void k_fluid_load(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ pos,
                             const double* __restrict__ vol, const T* __restrict__ F, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = idx[i];
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2], v = vol[i];
    const double fx = (double)F[3 * c], fy = (double)F[3 * c + 1], fz = (double)F[3 * c + 2];
    acc[0] += fx * v; acc[1] += fy * v; acc[2] += fz * v;
    acc[3] += (py * fz - pz * fy) * v; acc[4] += (pz * fx - px * fz) * v; acc[5] += (px * fy - py * fx) * v;
  }
  for (int k = 0; k < 6; ++k) {
    double r = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partials[k * kRedGrid + blockIdx.x] = r;
  }
}
template <typename T>
__global__ void k_diag_max(int64_t ncell, const T* __restrict__ div, const T* __restrict__ vc,
                           double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double a = 0, b = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    a = nanmax(a, fabs((double)div[c]));
    const double x = vc[3 * c], y = vc[3 * c + 1], z = vc[3 * c + 2];
    double ss = x * x + y * y;
    ss = ss + z * z;
    b = nanmax(b, sqrt(ss));
  }
  double ra = block_max(a, sh), rb = block_max(b, sh);
  if (threadIdx.x == 0) { partials[blockIdx.x] = ra; partials[kRedGrid + blockIdx.x] = rb; }
}

}  // namespace pyib

namespace {
using namespace pyib;

// ------------------------------------------------------- host launchers
template <typename T>
int flow_mat_t(const pyib_mesh* m, const pyib_work* ws, double rou, double visc, double hdt, const T* p, T* w,
               int mask_dir, double* dot_out, hipStream_t st) {
  T* face = static_cast<T*>(ws->face);
  T* ca = static_cast<T*>(ws->cell_a);
  T* cb = static_cast<T*>(ws->cell_b);
  k_curl<T><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->curl_ptr, m->curl_idx, p, face);
  k_interp<T><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
                                                     (const T*)m->rf1, (const T*)m->rf2, (const T*)m->vol_ci, face, ca);
  k_visc<T><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face, m->cf_other,
                                                   (const T*)m->aol, (const T*)m->vol_ci, (T)visc, (T)rou, (T)hdt, ca, cb);
  k_integr<T><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->f2c, (const T*)m->rf1, (const T*)m->rf2, cb, face);
  if (dot_out) {
    const int nb = red_grid_for(m->nedge);
    k_rot<T><<<nb, kBlock, 0, st>>>(m->nedge, m->rot_ptr, m->rot_idx, face, w, mask_dir ? m->dir_mask : nullptr, p,
                                    ws->partials);
    k_finalize<<<1, kFinBlock, 0, st>>>(ws->partials, nb, 1, dot_out);
  } else {
    k_rot<T><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, m->rot_ptr, m->rot_idx, face, w,
                                                    mask_dir ? m->dir_mask : nullptr, nullptr, nullptr);
  }
  return check_launch("pyib_flow_mat");
}

template <typename T>
int cg_iterate_t(const pyib_mesh* m, const pyib_work* ws, const pyib_cg* cg, double rou, double visc, double hdt,
                 int n_iter, int skip_last, hipStream_t st) {
  T* x = static_cast<T*>(cg->x);
  T* r = static_cast<T*>(cg->r);
  T* p = static_cast<T*>(cg->p);
  T* w = static_cast<T*>(cg->w);
  const T* prec = static_cast<const T*>(cg->prec);
  const int64_t n = m->nedge;
  const int nb = red_grid_for(n);
  T* face = static_cast<T*>(ws->face);
  T* ca = static_cast<T*>(ws->cell_a);
  T* cb = static_cast<T*>(ws->cell_b);
  for (int it = 0; it < n_iter; ++it) {
    k_curl<T><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->curl_ptr, m->curl_idx, p, face);
    k_interp<T><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
                                                       (const T*)m->rf1, (const T*)m->rf2, (const T*)m->vol_ci, face, ca);
    k_visc<T><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face, m->cf_other,
                                                     (const T*)m->aol, (const T*)m->vol_ci, (T)visc, (T)rou, (T)hdt, ca, cb);
    k_integr<T><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->f2c, (const T*)m->rf1, (const T*)m->rf2, cb, face);
    k_rot<T><<<nb, kBlock, 0, st>>>(n, m->rot_ptr, m->rot_idx, face, w, m->dir_mask, p, ws->partials);
    k_cg_alpha<<<1, kFinBlock, 0, st>>>(ws->partials, nb, cg->scal);
    k_cg_update<T><<<nb, kBlock, 0, st>>>(n, cg->scal, x, r, p, w, prec, ws->partials);
    k_cg_beta<<<1, kFinBlock, 0, st>>>(ws->partials, nb, cg->scal);
    if (!(skip_last && it == n_iter - 1)) {
      k_cg_dir<T><<<grid_for(n), kBlock, 0, st>>>(n, cg->scal, r, prec, p);
      k_cg_shift<<<1, 1, 0, st>>>(cg->scal);
    }
  }
  return check_launch("pyib_cg_iterate");
}

}  // namespace

// ------------------------------------------- allocation-free marker sort: device scratch
// One file-scope array per GPU the library runs on, held until the process ends.  Layout:
// key and value staging of up to kNoAllocMaxTaps entries each, then the radix sort's temporary
// storage.  The host never uses the array's name as an address: the address is fetched once per
// device with hipGetSymbolAddress (the name is only a host-side stand-in in host code).
constexpr int64_t kNoAllocMaxMarkers = PYIB_IB_NOALLOC_MAX_MARKERS;
constexpr int64_t kNoAllocMaxTaps = kNoAllocMaxMarkers * 27;
constexpr int64_t kNoAllocStage = ((kNoAllocMaxTaps * 4 + 255) / 256) * 256;   // bytes per staging array
constexpr int64_t kNoAllocBytes = PYIB_IB_NOALLOC_SCRATCH_BYTES;
static_assert(kNoAllocBytes > 2 * kNoAllocStage, "scratch too small for its staging arrays");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch`, looked up on the first call made on that device
int scratch_for_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return pyib::fail("pyib_ib_sort_noalloc: hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return pyib::fail_msg("pyib_ib_sort_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return pyib::fail("pyib_ib_sort_noalloc: hipGetSymbolAddress(scratch)", e);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

// ====================================================================== C ABI
extern "C" {
using namespace pyib;

const char* pyib_last_error(void) { return g_err; }

int pyib_device_count(int* count) {
  hipError_t e = hipGetDeviceCount(count);
  return e == hipSuccess ? 0 : fail("hipGetDeviceCount", e);
}

#define DISPATCH(dt, F64CALL, F32CALL) \
  ((dt) == PYIB_F64 ? (F64CALL) : (dt) == PYIB_F32 ? (F32CALL) : fail_msg("unknown dtype"))

int pyib_curl(const pyib_mesh* m, const void* edge, void* face, void* stream) {
  auto st = S(stream);
  if (m->dtype == PYIB_F64)
    k_curl<double><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->curl_ptr, m->curl_idx, (const double*)edge, (double*)face);
  else
    k_curl<float><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->curl_ptr, m->curl_idx, (const float*)edge, (float*)face);
  return check_launch("pyib_curl");
}

int pyib_rot(const pyib_mesh* m, const void* face, void* edge, void* stream) {
  auto st = S(stream);
  if (m->dtype == PYIB_F64)
    k_rot<double><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, m->rot_ptr, m->rot_idx, (const double*)face,
                                                         (double*)edge, nullptr, nullptr, nullptr);
  else
    k_rot<float><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, m->rot_ptr, m->rot_idx, (const float*)face,
                                                        (float*)edge, nullptr, nullptr, nullptr);
  return check_launch("pyib_rot");
}

int pyib_interp_f2c(const pyib_mesh* m, const void* face, void* cellvec, void* stream) {
  auto st = S(stream);
  if (m->dtype == PYIB_F64)
    k_interp<double><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
        (const double*)m->rf1, (const double*)m->rf2, (const double*)m->vol_ci, (const double*)face, (double*)cellvec);
  else
    k_interp<float><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
        (const float*)m->rf1, (const float*)m->rf2, (const float*)m->vol_ci, (const float*)face, (float*)cellvec);
  return check_launch("pyib_interp_f2c");
}

int pyib_integr_c2f(const pyib_mesh* m, const void* cellvec, void* face, void* stream) {
  auto st = S(stream);
  if (m->dtype == PYIB_F64)
    k_integr<double><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->f2c, (const double*)m->rf1,
        (const double*)m->rf2, (const double*)cellvec, (double*)face);
  else
    k_integr<float><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->f2c, (const float*)m->rf1,
        (const float*)m->rf2, (const float*)cellvec, (float*)face);
  return check_launch("pyib_integr_c2f");
}

int pyib_cdiv(const pyib_mesh* m, const void* face, void* cell, void* stream) {
  auto st = S(stream);
  if (m->dtype == PYIB_F64)
    k_cdiv<double><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
                                                          (const double*)face, nullptr, (double*)cell);
  else
    k_cdiv<float><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
                                                         (const float*)face, nullptr, (float*)cell);
  return check_launch("pyib_cdiv");
}

int pyib_flow_mat(const pyib_mesh* m, const pyib_work* ws, double rou, double visc, double hdt, const void* p,
                  void* w, int mask_dir, double* dot_out, void* stream) {
  return DISPATCH(m->dtype,
                  flow_mat_t<double>(m, ws, rou, visc, hdt, (const double*)p, (double*)w, mask_dir, dot_out, S(stream)),
                  flow_mat_t<float>(m, ws, rou, visc, hdt, (const float*)p, (float*)w, mask_dir, dot_out, S(stream)));
}

int pyib_calc_vel(const pyib_mesh* m, const void* s_in, void* s_out, void* U, void* vc, const double* lid,
                  void* stream) {
  auto st = S(stream);
  if (m->dtype != PYIB_F64) return fail_msg("pyib_calc_vel: float64 only");
  const double l0 = lid ? lid[0] : 0.0, l1 = lid ? lid[1] : 0.0, l2 = lid ? lid[2] : 0.0;
  if (s_in != s_out)
    k_copy_inlet<double><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, (const double*)s_in, (double*)s_out);
  if (m->ninlet > 0)
    k_set_inlet<double><<<grid_for(m->ninlet), kBlock, 0, st>>>(m->ninlet, m->inlet_edge,
                                                                (const double*)m->inlet_val, (double*)s_out);
  k_curl<double><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->curl_ptr, m->curl_idx, (const double*)s_out,
                                                        (double*)U);
  k_interp<double><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
      (const double*)m->rf1, (const double*)m->rf2, (const double*)m->vol_ci, (const double*)U, (double*)vc);
  if (m->nghost > 0)
    k_ghost<double><<<grid_for(m->nghost), kBlock, 0, st>>>(m->nghost, m->ncell, m->gh_face, m->gh_cell, m->gh_kind,
        (const double*)m->gh_nf, (const double*)m->gh_af2, (const double*)U, l0, l1, l2, (double*)vc);
  return check_launch("pyib_calc_vel");
}

int pyib_momentum_residual(const pyib_mesh* m, const void* U, const void* vc, double visc, const void* base,
                           double a_vc, double k_conv, double k_diff, const void* force, double k_force, void* out,
                           void* stream) {
  if (m->dtype != PYIB_F64) return fail_msg("pyib_momentum_residual: float64 only");
  k_momentum<double><<<grid_for(m->ncell), kBlock, 0, S(stream)>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
      m->cf_other, (const double*)m->aol, (const double*)m->vol_ci, (const double*)U, (const double*)vc, visc,
      (const double*)base, a_vc, k_conv, k_diff, (const double*)force, k_force, (double*)out);
  return check_launch("pyib_momentum_residual");
}

int pyib_curl_rhs(const pyib_mesh* m, const pyib_work* ws, const void* res_c, void* res_e, void* stream) {
  auto st = S(stream);
  if (m->dtype != PYIB_F64) return fail_msg("pyib_curl_rhs: float64 only");
  k_integr<double><<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->f2c, (const double*)m->rf1,
      (const double*)m->rf2, (const double*)res_c, (double*)ws->face);
  k_rot<double><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, m->rot_ptr, m->rot_idx, (const double*)ws->face,
                                                       (double*)res_e, nullptr, nullptr, nullptr);
  return check_launch("pyib_curl_rhs");
}

int pyib_cg_begin(const pyib_mesh* m, const pyib_work* ws, const pyib_cg* cg, const void* res_e, const void* s,
                  double tol1, double tol2, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(m->nedge);
  if (m->dtype == PYIB_F64)
    k_cg_begin<double><<<nb, kBlock, 0, st>>>(m->nedge, (const double*)res_e, (const double*)s, m->dir_mask,
        (const double*)cg->prec, (double*)cg->r, (double*)cg->x, (double*)cg->p, ws->partials);
  else
    k_cg_begin<float><<<nb, kBlock, 0, st>>>(m->nedge, (const float*)res_e, (const float*)s, m->dir_mask,
        (const float*)cg->prec, (float*)cg->r, (float*)cg->x, (float*)cg->p, ws->partials);
  k_cg_begin_fin<<<1, kFinBlock, 0, st>>>(ws->partials, nb, tol1, tol2, cg->scal);
  return check_launch("pyib_cg_begin");
}

int pyib_cg_iterate(const pyib_mesh* m, const pyib_work* ws, const pyib_cg* cg, double rou, double visc, double hdt,
                    int32_t n_iter, int32_t skip_last_direction, void* stream) {
  if (n_iter < 0) return fail_msg("pyib_cg_iterate: negative n_iter");
  return DISPATCH(m->dtype,
                  cg_iterate_t<double>(m, ws, cg, rou, visc, hdt, n_iter, skip_last_direction, S(stream)),
                  cg_iterate_t<float>(m, ws, cg, rou, visc, hdt, n_iter, skip_last_direction, S(stream)));
}

int pyib_cg_direction(const pyib_mesh* m, const pyib_work* ws, const pyib_cg* cg, void* stream) {
  (void)ws;
  auto st = S(stream);
  if (m->dtype == PYIB_F64)
    k_cg_dir<double><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, cg->scal, (const double*)cg->r,
                                                            (const double*)cg->prec, (double*)cg->p);
  else
    k_cg_dir<float><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, cg->scal, (const float*)cg->r,
                                                           (const float*)cg->prec, (float*)cg->p);
  k_cg_shift<<<1, 1, 0, st>>>(cg->scal);
  return check_launch("pyib_cg_direction");
}

int pyib_cg_true_residual(const pyib_mesh* m, const pyib_work* ws, const pyib_cg* cg, const void* res_e,
                          const void* s, void* tmp_edge, void* tr, double rou, double visc, double hdt, double* out2,
                          void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(m->nedge);
  int rc;
  if (m->dtype == PYIB_F64) {
    k_sub<double><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, (const double*)cg->x, (const double*)s,
                                                         (double*)tmp_edge);
    rc = flow_mat_t<double>(m, ws, rou, visc, hdt, (const double*)tmp_edge, (double*)tr, 0, nullptr, st);
    if (rc) return rc;
    k_true_res<double><<<nb, kBlock, 0, st>>>(m->nedge, (const double*)res_e, (const double*)tr, m->dir_mask,
                                              (const double*)cg->prec, (double*)tr, ws->partials);
  } else {
    k_sub<float><<<grid_for(m->nedge), kBlock, 0, st>>>(m->nedge, (const float*)cg->x, (const float*)s,
                                                        (float*)tmp_edge);
    rc = flow_mat_t<float>(m, ws, rou, visc, hdt, (const float*)tmp_edge, (float*)tr, 0, nullptr, st);
    if (rc) return rc;
    k_true_res<float><<<nb, kBlock, 0, st>>>(m->nedge, (const float*)res_e, (const float*)tr, m->dir_mask,
                                             (const float*)cg->prec, (float*)tr, ws->partials);
  }
  k_finalize<<<1, kFinBlock, 0, st>>>(ws->partials, nb, 2, out2);
  return check_launch("pyib_cg_true_residual");
}

int pyib_extrapolate(int64_t n, const double* s, const double* s_old, double dt, double dt_old, double* out,
                     void* stream) {
  k_extrapolate<<<grid_for(n), kBlock, 0, S(stream)>>>(n, s, s_old, dt, dt_old, out);
  return check_launch("pyib_extrapolate");
}

int pyib_convert(int dtype_in, const void* in, int dtype_out, void* out, int64_t n, void* stream) {
  auto st = S(stream);
  if (dtype_in == PYIB_F64 && dtype_out == PYIB_F32)
    k_convert<double, float><<<grid_for(n), kBlock, 0, st>>>(n, (const double*)in, (float*)out);
  else if (dtype_in == PYIB_F32 && dtype_out == PYIB_F64)
    k_convert<float, double><<<grid_for(n), kBlock, 0, st>>>(n, (const float*)in, (double*)out);
  else if (dtype_in == PYIB_F64 && dtype_out == PYIB_F64)
    k_convert<double, double><<<grid_for(n), kBlock, 0, st>>>(n, (const double*)in, (double*)out);
  else
    k_convert<float, float><<<grid_for(n), kBlock, 0, st>>>(n, (const float*)in, (float*)out);
  return check_launch("pyib_convert");
}

int pyib_add_into_f64(int dtype_a, const void* a, double* y, int64_t n, void* stream) {
  auto st = S(stream);
  if (dtype_a == PYIB_F64)
    k_add_into<double><<<grid_for(n), kBlock, 0, st>>>(n, (const double*)a, y);
  else
    k_add_into<float><<<grid_for(n), kBlock, 0, st>>>(n, (const float*)a, y);
  return check_launch("pyib_add_into_f64");
}

int pyib_dot(int dtype, const void* a, const void* b, const void* c, int64_t n, double* partials, double* out,
             void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(n);
  if (dtype == PYIB_F64)
    k_dot<double><<<nb, kBlock, 0, st>>>(n, (const double*)a, (const double*)b, (const double*)c, partials);
  else
    k_dot<float><<<nb, kBlock, 0, st>>>(n, (const float*)a, (const float*)b, (const float*)c, partials);
  k_finalize<<<1, kFinBlock, 0, st>>>(partials, nb, 1, out);
  return check_launch("pyib_dot");
}

int pyib_axpby(int dtype, double a, const void* x, double b, const void* y, void* out, int64_t n, void* stream) {
  auto st = S(stream);
  if (dtype == PYIB_F64)
    k_axpby<double><<<grid_for(n), kBlock, 0, st>>>(n, a, (const double*)x, b, (const double*)y, (double*)out);
  else
    k_axpby<float><<<grid_for(n), kBlock, 0, st>>>(n, (float)a, (const float*)x, (float)b, (const float*)y, (float*)out);
  return check_launch("pyib_axpby");
}

int pyib_ib_stencil(int64_t nmark, const double* markers, double pr0, double pr1, double pr2, double dx, double hh,
                    int32_t nx, int32_t ny, int32_t nz, int32_t* cell, double* w, void* stream) {
  k_ib_stencil<<<grid_for(nmark * 27), kBlock, 0, S(stream)>>>(nmark, markers, pr0, pr1, pr2, dx, hh, nx, ny, nz,
                                                              cell, w);
  return check_launch("pyib_ib_stencil");
}

int pyib_ib_interp(int dtype, int64_t nmark, const int32_t* cell, const double* w, double hh3, const void* vc,
                   double* vel, void* stream) {
  auto st = S(stream);
  if (dtype == PYIB_F64)
    k_ib_interp<double><<<grid_for(nmark), kBlock, 0, st>>>(nmark, cell, w, hh3, (const double*)vc, vel);
  else
    k_ib_interp<float><<<grid_for(nmark), kBlock, 0, st>>>(nmark, cell, w, hh3, (const float*)vc, vel);
  return check_launch("pyib_ib_interp");
}

int pyib_ib_sort_temp_bytes(int64_t n, int64_t* bytes) {
  size_t b = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const int32_t*)nullptr, (int32_t*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, (int)n);
  if (e != hipSuccess) return fail("pyib_ib_sort_temp_bytes", e);
  *bytes = (int64_t)b;
  return 0;
}

int pyib_ib_sort(int64_t nmark, const int32_t* cell, int32_t* keys_tmp, int32_t* vals_tmp, int32_t* keys_sorted,
                 int32_t* vals_sorted, void* temp, int64_t temp_bytes, void* stream) {
  const int64_t n = nmark * 27;
  auto st = S(stream);
  k_ib_keys<<<grid_for(n), kBlock, 0, st>>>(n, cell, keys_tmp, vals_tmp);
  size_t b = (size_t)temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, b, (const int32_t*)keys_tmp, keys_sorted,
                                                    (const int32_t*)vals_tmp, vals_sorted, (int)n, 0,
                                                    (int)(sizeof(int32_t) * 8), st);
  if (e != hipSuccess) return fail("pyib_ib_sort", e);
  return check_launch("pyib_ib_sort");
}

int pyib_ib_sort_noalloc(int64_t nmark, const int32_t* cell, int32_t* keys_sorted, int32_t* vals_sorted,
                         void* stream) {
  if (nmark < 0 || nmark > kNoAllocMaxMarkers) {
    std::snprintf(g_err, sizeof(g_err), "pyib_ib_sort_noalloc: %lld markers outside [0, %lld] (the static scratch's "
                  "limit; use pyib_ib_sort with caller workspaces)", (long long)nmark, (long long)kNoAllocMaxMarkers);
    return 2;
  }
  if (nmark == 0) return 0;
  void* base = nullptr;
  if (int rc = scratch_for_current_device(&base)) return rc;
  const int64_t n = nmark * 27;
  size_t need = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, (const int32_t*)nullptr, (int32_t*)nullptr,
                                                    (const int32_t*)nullptr, (int32_t*)nullptr, (int)n);
  if (e != hipSuccess) return fail("pyib_ib_sort_noalloc", e);
  const size_t avail = (size_t)(kNoAllocBytes - 2 * kNoAllocStage);
  if (need > avail) {
    std::snprintf(g_err, sizeof(g_err), "pyib_ib_sort_noalloc: radix-sort storage %zu bytes exceeds the %zu in the "
                  "static scratch", need, avail);
    return 2;
  }
  unsigned char* b = static_cast<unsigned char*>(base);
  int32_t* keys_tmp = reinterpret_cast<int32_t*>(b);
  int32_t* vals_tmp = reinterpret_cast<int32_t*>(b + kNoAllocStage);
  void* temp = b + 2 * kNoAllocStage;
  return pyib_ib_sort(nmark, cell, keys_tmp, vals_tmp, keys_sorted, vals_sorted, temp, (int64_t)avail, stream);
}

int pyib_ib_spread(int dtype, int64_t nmark, const int32_t* keys_sorted, const int32_t* vals_sorted, const double* w,
                   const double* flag, const double* ds3, double scale, int accumulate, void* out, void* stream) {
  const int64_t n = nmark * 27;
  auto st = S(stream);
  if (dtype == PYIB_F64)
    k_ib_spread<double><<<grid_for(n), kBlock, 0, st>>>(n, keys_sorted, vals_sorted, w, flag, ds3, scale, accumulate,
                                                        (double*)out);
  else
    k_ib_spread<float><<<grid_for(n), kBlock, 0, st>>>(n, keys_sorted, vals_sorted, w, flag, ds3, scale, accumulate,
                                                       (float*)out);
  return check_launch("pyib_ib_spread");
}

int pyib_ib_force(int64_t nmark, const double* vel_desired, const double* vel_ib, double dt, double* f_lag,
                  void* stream) {
  k_ib_force<<<grid_for(3 * nmark), kBlock, 0, S(stream)>>>(3 * nmark, vel_desired, vel_ib, dt, f_lag);
  return check_launch("pyib_ib_force");
}

int pyib_interior_momenta(int dtype, int64_t n, const int32_t* idx, const double* pos, const double* vol,
                          const void* vc, double cx, double cy, double cz, double radius, double width, double rho,
                          int fixed, const double* fixed_wv, double* partials, double* out7, void* stream) {
  (void)rho;
  auto st = S(stream);
  const int nb = red_grid_for(n);
  if (dtype == PYIB_F64)
    k_interior<double><<<nb, kBlock, 0, st>>>(n, idx, pos, vol, (const double*)vc, cx, cy, cz, radius, width, fixed,
                                              fixed_wv, partials);
  else
    k_interior<float><<<nb, kBlock, 0, st>>>(n, idx, pos, vol, (const float*)vc, cx, cy, cz, radius, width, fixed,
                                             fixed_wv, partials);
  k_finalize<<<1, kFinBlock, 0, st>>>(partials, nb, 7, out7);
  return check_launch("pyib_interior_momenta");
}

int pyib_fluid_load(int dtype, int64_t n, const int32_t* idx, const double* pos, const double* vol, const void* F,
                    double* partials, double* out6, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(n);
  if (dtype == PYIB_F64)
    k_fluid_load<double><<<nb, kBlock, 0, st>>>(n, idx, pos, vol, (const double*)F, partials);
  else
    k_fluid_load<float><<<nb, kBlock, 0, st>>>(n, idx, pos, vol, (const float*)F, partials);
  k_finalize<<<1, kFinBlock, 0, st>>>(partials, nb, 6, out6);
  return check_launch("pyib_fluid_load");
}

int pyib_diag_max(const pyib_mesh* m, const void* U, const void* vc, void* tmp_cell, double* out2, void* stream) {
  auto st = S(stream);
  if (m->dtype != PYIB_F64) return fail_msg("pyib_diag_max: float64 only");
  k_cdiv<double><<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
                                                        (const double*)U, (const double*)m->vol_ci, (double*)tmp_cell);
  // partials live after the cell temp (caller sizes tmp_cell >= ncell + 2*PYIB_MAX_PARTIALS)
  double* partials = (double*)tmp_cell + m->ncell;
  const int nb = red_grid_for(m->ncell);
  k_diag_max<double><<<nb, kBlock, 0, st>>>(m->ncell, (const double*)tmp_cell, (const double*)vc, partials);
  k_finalize_max<<<1, kFinBlock, 0, st>>>(partials, nb, 2, out2);
  return check_launch("pyib_diag_max");
}

}  // extern "C"
