// adjoint.hip -- transposes of the time-step operators (gfx950, float64).
//
// Each kernel is the exact transpose of one forward operator of the step
// (ib_driver_np.py:267-303, 458-527; ib_immersed.py:63-120; ib_interior.py:57-70)
// written as a gather over the same incidence tables, so the backward pass
// is deterministic and needs no atomics.  The elliptic operator is symmetric
// (A = C^T I V P C with P = diag(vol_ci) I^T), so the adjoint of a Helmholtz
// solve is the same PCG with the adjoint right-hand side.
#include "common.h"

namespace pyib {
namespace {

// ---- calc_vel^T ---------------------------------------------------------
// (1) ghost rows of g_vc fold into their owner cells (cell-centric gather over
//     the cell's boundary faces), kinds as in k_ghost.
__global__ void k_ghost_adj_cells(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                                  const int32_t* __restrict__ other, const int32_t* __restrict__ gkind,
                                  const double* __restrict__ gnf, const double* __restrict__ gaf2,
                                  double* __restrict__ g) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int j = ptr[c]; j < mid[c]; ++j) {   // boundary faces are always first-owner entries
      const int o = other[j];
      if (o < ncell) continue;
      const int64_t k = o - ncell;
      const double q0 = g[3 * (int64_t)o], q1 = g[3 * (int64_t)o + 1], q2 = g[3 * (int64_t)o + 2];
      switch (gkind[k]) {
        case 0: break;                                     // inlet: ghost independent of vc
        case 1: {                                          // v1 + ((U - v1.n) n)/Af^2
          const double n0 = gnf[3 * k], n1 = gnf[3 * k + 1], n2 = gnf[3 * k + 2];
          const double d = (q0 * n0 + q1 * n1 + q2 * n2) / gaf2[k];
          a0 += q0 - d * n0; a1 += q1 - d * n1; a2 += q2 - d * n2;
          break;
        }
        case 2: case 3: a0 -= q0; a1 -= q1; a2 -= q2; break;   // wall / lid: -v1 (+ constant)
        default: a0 += q0; a1 += q1; a2 += q2; break;           // spanwise copy
      }
    }
    g[3 * c] += a0; g[3 * c + 1] += a1; g[3 * c + 2] += a2;
  }
}

// (2) g_U[f] = interp^T g  (+ the ghost's dependence on U) + g_U_in[f]
__global__ void k_interp_adj(int64_t nface, int64_t ncell, const int32_t* __restrict__ f2c,
                             const double* __restrict__ rf1, const double* __restrict__ rf2,
                             const double* __restrict__ vol_ci, const int32_t* __restrict__ face_ghost,
                             const int32_t* __restrict__ gkind, const double* __restrict__ gnf,
                             const double* __restrict__ gaf2, const double* __restrict__ g,
                             const double* __restrict__ gU_in, double* __restrict__ gU) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nface; f += (int64_t)gridDim.x * blockDim.x) {
    const int c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
    double v = vol_ci[c1] * (g[3 * c1] * rf1[3 * f] + g[3 * c1 + 1] * rf1[3 * f + 1] + g[3 * c1 + 2] * rf1[3 * f + 2]);
    if (c2 >= 0)
      v -= vol_ci[c2] * (g[3 * c2] * rf2[3 * f] + g[3 * c2 + 1] * rf2[3 * f + 1] + g[3 * c2 + 2] * rf2[3 * f + 2]);
    const int k = face_ghost[f];
    if (k >= 0 && gkind[k] <= 1) {
      const int64_t o = ncell + k;
      v += (g[3 * o] * gnf[3 * k] + g[3 * o + 1] * gnf[3 * k + 1] + g[3 * o + 2] * gnf[3 * k + 2]) / gaf2[k];
    }
    if (gU_in) v += gU_in[f];
    gU[f] = v;
  }
}

__global__ void k_zero_idx(int64_t n, const int32_t* __restrict__ idx, double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[idx[i]] = 0.0;
}

// integr_c2f^T (ib_core_np.py:144-148 transposed): g_v[c] = sum_{faces with c = c1} g_F rf1
// - sum_{faces with c = c2} g_F rf2, a cell-centric gather over the cell's face list.
__global__ void k_integr_adj(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                             const int32_t* __restrict__ face, const double* __restrict__ rf1,
                             const double* __restrict__ rf2, const double* __restrict__ gF, double* __restrict__ gv) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0;
    const int j0 = ptr[c], jm = mid[c], j1 = ptr[c + 1];
    for (int j = j0; j < jm; ++j) {
      const int f = face[j];
      const double u = gF[f];
      a0 += u * rf1[3 * f]; a1 += u * rf1[3 * f + 1]; a2 += u * rf1[3 * f + 2];
    }
    for (int j = jm; j < j1; ++j) {
      const int f = face[j];
      const double u = gF[f];
      b0 += u * rf2[3 * f]; b1 += u * rf2[3 * f + 1]; b2 += u * rf2[3 * f + 2];
    }
    gv[3 * c] = a0 - b0; gv[3 * c + 1] = a1 - b1; gv[3 * c + 2] = a2 - b2;
  }
}

// ---- momentum residual^T ---------------------------------------------
// out = [base - a*vc | a*vc] - kc*conv(U, vc) + kd*diff(vc) (+ kf*F); given g_out:
// g_vc_ext += ..., g_U += ...   (cell, ghost and face passes)
__global__ void k_mom_adj_cells(int64_t ncell, const int32_t* __restrict__ ptr, const int32_t* __restrict__ mid,
                                const int32_t* __restrict__ face, const int32_t* __restrict__ other,
                                const double* __restrict__ aol, const double* __restrict__ vol_ci,
                                const double* __restrict__ U, double visc, double own_a, double kc, double kd,
                                const double* __restrict__ gout, double* __restrict__ gvc) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < ncell; x += (int64_t)gridDim.x * blockDim.x) {
    const double vx = vol_ci[x];
    double s1 = 0, sa = 0;
    double n0 = 0, n1 = 0, n2 = 0;
    for (int j = ptr[x]; j < ptr[x + 1]; ++j) {
      const int f = face[j], o = other[j];
      const double sg = j < mid[x] ? 1.0 : -1.0;
      const double wf = 0.5 * U[f];
      s1 += sg * wf;
      sa += aol[f];
      if (o < ncell) {       // x as the opposite owner of o's face: sigma_o = -sg
        const double vo = vol_ci[o];
        const double cc = -kc * vo * (-sg) * wf;      // g_conv[o] = -kc g_out[o]
        const double cd = kd * vo * visc * aol[f];    // g_diff[o] = kd g_out[o]
        const double k = cc + cd;
        n0 += k * gout[3 * o]; n1 += k * gout[3 * o + 1]; n2 += k * gout[3 * o + 2];
      }
    }
    const double own = own_a + (-kc) * vx * s1 + kd * (-vx * visc * sa);
    gvc[3 * x] += own * gout[3 * x] + n0;
    gvc[3 * x + 1] += own * gout[3 * x + 1] + n1;
    gvc[3 * x + 2] += own * gout[3 * x + 2] + n2;
  }
}
__global__ void k_mom_adj_ghosts(int64_t nghost, int64_t ncell, const int32_t* __restrict__ gface,
                                 const int32_t* __restrict__ gcell, const double* __restrict__ aol,
                                 const double* __restrict__ vol_ci, const double* __restrict__ U, double visc,
                                 double kc, double kd, const double* __restrict__ gout, double* __restrict__ gvc) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nghost; k += (int64_t)gridDim.x * blockDim.x) {
    const int f = gface[k], c = gcell[k];
    const double vo = vol_ci[c];
    const double kk = -kc * vo * (0.5 * U[f]) + kd * vo * visc * aol[f];
    const int64_t g = ncell + k;
    gvc[3 * g] += kk * gout[3 * c]; gvc[3 * g + 1] += kk * gout[3 * c + 1]; gvc[3 * g + 2] += kk * gout[3 * c + 2];
  }
}
__global__ void k_mom_adj_faces(int64_t nface, int64_t ncell, const int32_t* __restrict__ f2c,
                                const int32_t* __restrict__ face_ghost, const double* __restrict__ vol_ci,
                                const double* __restrict__ vc, double kc, const double* __restrict__ gout,
                                double* __restrict__ gU) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nface; f += (int64_t)gridDim.x * blockDim.x) {
    const int c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
    const int64_t o1 = c2 >= 0 ? (int64_t)c2 : ncell + face_ghost[f];
    // conv[c1] += vol_ci[c1] * 0.5 U (vc[c1] + vc[o1]);  conv[c2] -= vol_ci[c2] * 0.5 U (vc[c2] + vc[c1])
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s += gout[3 * c1 + d] * (vc[3 * c1 + d] + vc[3 * o1 + d]);
    double v = (-kc) * vol_ci[c1] * 0.5 * s;
    if (c2 >= 0) {
      double t = 0.0;
      for (int d = 0; d < 3; ++d) t += gout[3 * c2 + d] * (vc[3 * c2 + d] + vc[3 * c1 + d]);
      v -= (-kc) * vol_ci[c2] * 0.5 * t;
    }
    gU[f] += v;
  }
}

// ---- immersed boundary -------------------------------------------------
__device__ __forceinline__ double delta_h3(double r, double h) {
  const double a = fabs(r / h);
  if (a <= 0.5) return (1.0 + sqrt(fmax(1.0 - 3.0 * a * a, 0.0))) / (3.0 * h);
  if (a <= 1.5) {
    const double q = 1.0 - a;
    return (5.0 - 3.0 * a - sqrt(fmax(1.0 - 3.0 * (q * q), 0.0))) / (6.0 * h);
  }
  return 0.0;
}
// d delta / d r: derivative of the two branches of delta_h3 (the kernel is C^1 at a = 0.5 and 1.5)
__device__ __forceinline__ double ddelta_h3(double r, double h) {
  const double a = fabs(r / h);
  const double sg = r >= 0.0 ? 1.0 : -1.0;
  if (a <= 0.5) {
    const double t = sqrt(fmax(1.0 - 3.0 * a * a, 1e-300));
    return sg * (-a / t) / (h * h);
  }
  if (a <= 1.5) {
    const double q = 1.0 - a;
    const double t = sqrt(fmax(1.0 - 3.0 * q * q, 1e-300));
    return sg * (-3.0 - 3.0 * q / t) / (6.0 * h * h);
  }
  return 0.0;
}

// dw/dm for every (marker, tap): w = d(gx) d(gy) d(gz), g = grid - m  => dw/dm = -d'(g)...
__global__ void k_ib_dweights(int64_t nmark, const double* __restrict__ mk, double pr0, double pr1, double pr2,
                              double dx, double hh, int nx, int ny, int nz, double* __restrict__ dw) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nmark * 27; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = i / 27;
    const int t = (int)(i % 27);
    const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
    const double mx = mk[3 * m], my = mk[3 * m + 1], mz = mk[3 * m + 2];
    const long ii = (long)rint((mx - pr0) / dx) + di;
    const long jj = (long)rint((my - pr1) / dx) + dj;
    const long kk = (long)rint((mz - pr2) / dx) + dk;
    const bool inb = ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz;
    if (!inb) { dw[3 * i] = dw[3 * i + 1] = dw[3 * i + 2] = 0.0; continue; }
    const double gx = (pr0 + (double)ii * dx) - mx;
    const double gy = (pr1 + (double)jj * dx) - my;
    const double gz = (pr2 + (double)kk * dx) - mz;
    const double ax = delta_h3(gx, hh), ay = delta_h3(gy, hh), az = delta_h3(gz, hh);
    dw[3 * i] = -ddelta_h3(gx, hh) * ay * az;
    dw[3 * i + 1] = -ax * ddelta_h3(gy, hh) * az;
    dw[3 * i + 2] = -ax * ay * ddelta_h3(gz, hh);
  }
}

// out[m] += scale[m] * sum_t (field[cell] . vec[m]) * dw[m,t]   (marker-position gradient)
__global__ void k_ib_marker_grad(int64_t nmark, const int32_t* __restrict__ cell, const double* __restrict__ dw,
                                 const double* __restrict__ field, const double* __restrict__ vec,
                                 const double* __restrict__ scale, double scale_all, double* __restrict__ out) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < nmark; m += (int64_t)gridDim.x * blockDim.x) {
    double a0 = 0, a1 = 0, a2 = 0;
    const double v0 = vec[3 * m], v1 = vec[3 * m + 1], v2 = vec[3 * m + 2];
    for (int t = 0; t < 27; ++t) {
      const int c = cell[27 * m + t];
      if (c < 0) continue;
      const double q = field[3 * (int64_t)c] * v0 + field[3 * (int64_t)c + 1] * v1 + field[3 * (int64_t)c + 2] * v2;
      const int64_t i = 27 * m + t;
      a0 += q * dw[3 * i]; a1 += q * dw[3 * i + 1]; a2 += q * dw[3 * i + 2];
    }
    const double s = (scale ? scale[m] : 1.0) * scale_all;
    out[3 * m] += s * a0; out[3 * m + 1] += s * a1; out[3 * m + 2] += s * a2;
  }
}

// out[m] = scale[m] * sum_t field[cell] w[m,t]   (spread^T with the marker volume)
__global__ void k_ib_gather(int64_t nmark, const int32_t* __restrict__ cell, const double* __restrict__ w,
                            const double* __restrict__ field, const double* __restrict__ scale,
                            double* __restrict__ out) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < nmark; m += (int64_t)gridDim.x * blockDim.x) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int t = 0; t < 27; ++t) {
      const int c = cell[27 * m + t];
      if (c < 0) continue;
      const double wt = w[27 * m + t];
      a0 += field[3 * (int64_t)c] * wt; a1 += field[3 * (int64_t)c + 1] * wt; a2 += field[3 * (int64_t)c + 2] * wt;
    }
    const double s = scale[m];
    out[3 * m] = s * a0; out[3 * m + 1] = s * a1; out[3 * m + 2] = s * a2;
  }
}

// ---- loads -------------------------------------------------------------
// g_F[idx[i]] += gvec . vol[i] (force_fluid = sum F*vol over the support cells)
__global__ void k_fluid_load_adj(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ vol,
                                 double g0, double g1, double g2, double* __restrict__ gF) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx[i];
    gF[3 * c] += g0 * vol[i]; gF[3 * c + 1] += g1 * vol[i]; gF[3 * c + 2] += g2 * vol[i];
  }
}
// linear momentum P = rho sum v wv(center): g_v[idx] += rho gP wv ; partials of dP/dcenter . gP
__global__ void k_momentum_adj(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ pos,
                               const double* __restrict__ vol, const double* __restrict__ vc, double cx, double cy,
                               double cz, double radius, double width, double rho, double g0, double g1, double g2,
                               double* __restrict__ gvc, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double acc[3] = {0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double px = pos[3 * i] - cx, py = pos[3 * i + 1] - cy, pz = pos[3 * i + 2] - cz;
    const double d = sqrt(px * px + py * py + pz * pz);
    const double raw = ((radius + 0.5 * width) - d) / width;
    const double ind = fmin(fmax(raw, 0.0), 1.0);
    const double wv = ind * vol[i];
    const int64_t c = idx[i];
    gvc[3 * c] += rho * g0 * wv; gvc[3 * c + 1] += rho * g1 * wv; gvc[3 * c + 2] += rho * g2 * wv;
    if (raw > 0.0 && raw < 1.0 && d > 0.0) {
      // d wv / d center = vol/width * (pos - center)/d
      const double vg = rho * (vc[3 * c] * g0 + vc[3 * c + 1] * g1 + vc[3 * c + 2] * g2) * vol[i] / (width * d);
      acc[0] += vg * px; acc[1] += vg * py; acc[2] += vg * pz;
    }
  }
  for (int k = 0; k < 3; ++k) {
    double r = block_sum(acc[k], sh);
    if (threadIdx.x == 0) partials[k * kRedGrid + blockIdx.x] = r;
  }
}

}  // namespace
}  // namespace pyib

extern "C" {
using namespace pyib;

int pyib_calc_vel_adjoint(const pyib_mesh* m, const int32_t* face_ghost, double* g_vc, const double* g_U_in,
                          double* g_U, double* g_s, void* stream) {
  auto st = S(stream);
  if (m->dtype != PYIB_F64) return fail_msg("pyib_calc_vel_adjoint: float64 only");
  k_ghost_adj_cells<<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_other, m->gh_kind,
                                                          (const double*)m->gh_nf, (const double*)m->gh_af2, g_vc);
  k_interp_adj<<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->ncell, m->f2c, (const double*)m->rf1,
      (const double*)m->rf2, (const double*)m->vol_ci, face_ghost, m->gh_kind, (const double*)m->gh_nf,
      (const double*)m->gh_af2, g_vc, g_U_in, g_U);
  int rc = check_launch("pyib_calc_vel_adjoint");
  if (rc) return rc;
  rc = pyib_rot(m, g_U, g_s, stream);          // curl^T
  if (rc) return rc;
  if (m->ninlet > 0) k_zero_idx<<<grid_for(m->ninlet), kBlock, 0, st>>>(m->ninlet, m->inlet_edge, g_s);
  return check_launch("pyib_calc_vel_adjoint");
}

int pyib_curl_rhs_adjoint(const pyib_mesh* m, const pyib_work* ws, const double* g_edge, double* g_res_c,
                          void* stream) {
  if (m->dtype != PYIB_F64) return fail_msg("pyib_curl_rhs_adjoint: float64 only");
  int rc = pyib_curl(m, g_edge, ws->face, stream);     // rot^T = curl
  if (rc) return rc;
  k_integr_adj<<<grid_for(m->ncell), kBlock, 0, S(stream)>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face,
      (const double*)m->rf1, (const double*)m->rf2, (const double*)ws->face, g_res_c);
  return check_launch("pyib_curl_rhs_adjoint");
}

int pyib_momentum_residual_adjoint(const pyib_mesh* m, const int32_t* face_ghost, const double* U, const double* vc,
                                   double visc, int has_base, double a_vc, double k_conv, double k_diff,
                                   const double* g_out, double* g_vc, double* g_U, void* stream) {
  auto st = S(stream);
  if (m->dtype != PYIB_F64) return fail_msg("pyib_momentum_residual_adjoint: float64 only");
  const double own_a = has_base ? -a_vc : a_vc;
  k_mom_adj_cells<<<grid_for(m->ncell), kBlock, 0, st>>>(m->ncell, m->cf_ptr, m->cf_mid, m->cf_face, m->cf_other,
      (const double*)m->aol, (const double*)m->vol_ci, U, visc, own_a, k_conv, k_diff, g_out, g_vc);
  if (m->nghost > 0)
    k_mom_adj_ghosts<<<grid_for(m->nghost), kBlock, 0, st>>>(m->nghost, m->ncell, m->gh_face, m->gh_cell,
        (const double*)m->aol, (const double*)m->vol_ci, U, visc, k_conv, k_diff, g_out, g_vc);
  k_mom_adj_faces<<<grid_for(m->nface), kBlock, 0, st>>>(m->nface, m->ncell, m->f2c, face_ghost,
      (const double*)m->vol_ci, vc, k_conv, g_out, g_U);
  return check_launch("pyib_momentum_residual_adjoint");
}

int pyib_ib_dweights(int64_t nmark, const double* markers, double pr0, double pr1, double pr2, double dx, double hh,
                     int32_t nx, int32_t ny, int32_t nz, double* dw, void* stream) {
  k_ib_dweights<<<grid_for(nmark * 27), kBlock, 0, S(stream)>>>(nmark, markers, pr0, pr1, pr2, dx, hh, nx, ny, nz,
                                                               dw);
  return check_launch("pyib_ib_dweights");
}

int pyib_ib_marker_grad(int64_t nmark, const int32_t* cell, const double* dw, const double* field, const double* vec,
                        const double* scale, double scale_all, double* out, void* stream) {
  k_ib_marker_grad<<<grid_for(nmark), kBlock, 0, S(stream)>>>(nmark, cell, dw, field, vec, scale, scale_all, out);
  return check_launch("pyib_ib_marker_grad");
}

int pyib_ib_gather(int64_t nmark, const int32_t* cell, const double* w, const double* field, const double* scale,
                   double* out, void* stream) {
  k_ib_gather<<<grid_for(nmark), kBlock, 0, S(stream)>>>(nmark, cell, w, field, scale, out);
  return check_launch("pyib_ib_gather");
}

int pyib_fluid_load_adjoint(int64_t n, const int32_t* idx, const double* vol, double g0, double g1, double g2,
                            double* gF, void* stream) {
  k_fluid_load_adj<<<grid_for(n), kBlock, 0, S(stream)>>>(n, idx, vol, g0, g1, g2, gF);
  return check_launch("pyib_fluid_load_adjoint");
}

int pyib_interior_momentum_adjoint(int64_t n, const int32_t* idx, const double* pos, const double* vol,
                                   const double* vc, double cx, double cy, double cz, double radius, double width,
                                   double rho, double g0, double g1, double g2, double* gvc, double* partials,
                                   double* gcenter3, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(n);
  k_momentum_adj<<<nb, kBlock, 0, st>>>(n, idx, pos, vol, vc, cx, cy, cz, radius, width, rho, g0, g1, g2, gvc,
                                        partials);
  k_finalize<<<1, kFinBlock, 0, st>>>(partials, nb, 3, gcenter3);
  return check_launch("pyib_interior_momentum_adjoint");
}

}  // extern "C"
