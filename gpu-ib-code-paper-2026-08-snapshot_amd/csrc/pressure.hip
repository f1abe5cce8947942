// pressure.hip -- offline pressure recovery of a saved field on the GPU.
//
// The reference recovers the cell pressure whose face jumps best fit the saved pressure gradient
// (src/pyib/solver/ib_postprocess.py:49-105): on the cell graph, with w_f = A_f / |x2 - x1| and
// j_f = (g1 + g2)/2 . (x2 - x1) on every internal face, it solves the weighted graph Laplacian
// L p = I^T W j with cell 0 pinned by Jacobi-preconditioned CG, then shifts p so that the
// boundary-adjacent cells have zero mean.  Here the same system is solved matrix-free: each cell
// gathers its internal faces from a cell-major adjacency (built on the host by a counting sort,
// faces in face order), so there are no atomics and every reduction is a fixed two-stage sum.
// The CG scalars stay on the device; the host reads |r|^2 at its own cadence.
#include <new>

#include "common.h"

namespace pyib {
namespace {

enum { P_RHO = 0, P_PQ = 1, P_ALPHA = 2, P_BETA = 3, P_RR = 4, P_BB = 5 };

// per face: weight and jump (0 on boundary faces)
__global__ void k_pg_face(int64_t m, const int32_t* __restrict__ fc, const double* __restrict__ pos,
                          const double* __restrict__ grad, const double* __restrict__ area, double* __restrict__ w,
                          double* __restrict__ jmp) {
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < m; f += (int64_t)gridDim.x * blockDim.x) {
    const int c1 = fc[2 * f], c2 = fc[2 * f + 1];
    if (c2 < 0) {
      w[f] = 0.0;
      jmp[f] = 0.0;
      continue;
    }
    const double d0 = pos[3 * c2] - pos[3 * c1], d1 = pos[3 * c2 + 1] - pos[3 * c1 + 1],
                 d2 = pos[3 * c2 + 2] - pos[3 * c1 + 2];
    const double dist = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    const double j = 0.5 * (((grad[3 * c1] + grad[3 * c2]) * d0 + (grad[3 * c1 + 1] + grad[3 * c2 + 1]) * d1) +
                            (grad[3 * c1 + 2] + grad[3 * c2 + 2]) * d2);
    w[f] = area[f] / dist;
    jmp[f] = j;
  }
}

// per cell: right-hand side (I^T W j), inverse diagonal; cell 0 is pinned (row and unknown 0);
// CG start x = 0, r = b, z = D^-1 r, p = z; partials of r.z and r.r
__global__ void k_pg_begin(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ sface,
                           const double* __restrict__ w, const double* __restrict__ jmp, double* __restrict__ dinv,
                           double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                           double* __restrict__ p, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double rz = 0.0, rr = 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    double b = 0.0, dg = 0.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
      const int s = sface[k];
      const int f = s >= 0 ? s : -s - 1;
      const double wj = w[f] * jmp[f];
      b += s >= 0 ? wj : -wj;
      dg += w[f];
    }
    const double di = (c == 0 || dg == 0.0) ? 0.0 : 1.0 / dg;
    const double rc = c == 0 ? 0.0 : b;
    dinv[c] = di;
    x[c] = 0.0;
    r[c] = rc;
    z[c] = di * rc;
    p[c] = di * rc;
    rz += rc * (di * rc);
    rr += rc * rc;
  }
  const double a = block_sum(rz, sh), bsum = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a;
    partials[kRedGrid + blockIdx.x] = bsum;
  }
}

// q = L p on the pinned system (q[0] = 0, p[0] stays 0); partials of p.q
__global__ void k_pg_apply(int64_t n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ nbr,
                           const int32_t* __restrict__ sface, const double* __restrict__ w,
                           const double* __restrict__ p, double* __restrict__ q, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double pq = 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    if (c != 0) {
      const double pc = p[c];
      for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
        const int s = sface[k];
        const int f = s >= 0 ? s : -s - 1;
        const int o = nbr[k];
        acc += w[f] * (pc - (o == 0 ? 0.0 : p[o]));
      }
      pq += pc * acc;
    }
    q[c] = acc;
  }
  const double a = block_sum(pq, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

// alpha = rho / p.q (0 if p.q is 0)
__global__ void k_pg_alpha(const double* __restrict__ partials, int nb, double* __restrict__ scal) {
  __shared__ double sh[kFinBlock / 64];
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += partials[b];
  const double pq = block_sum(v, sh);
  if (threadIdx.x == 0) {
    scal[P_PQ] = pq;
    scal[P_ALPHA] = pq != 0.0 ? scal[P_RHO] / pq : 0.0;
  }
}

// x += alpha p, r -= alpha q, z = D^-1 r; partials of r.z and r.r
__global__ void k_pg_update(int64_t n, const double* __restrict__ scal, double* __restrict__ x,
                            double* __restrict__ r, const double* __restrict__ p, const double* __restrict__ q,
                            const double* __restrict__ dinv, double* __restrict__ z, double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  const double alpha = scal[P_ALPHA];
  double rz = 0.0, rr = 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    x[c] += alpha * p[c];
    const double rc = r[c] - alpha * q[c];
    const double zc = dinv[c] * rc;
    r[c] = rc;
    z[c] = zc;
    rz += rc * zc;
    rr += rc * rc;
  }
  const double a = block_sum(rz, sh), b = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a;
    partials[kRedGrid + blockIdx.x] = b;
  }
}

// rho_new = r.z, rr = r.r, beta = rho_new / rho; at the start (first != 0) also |b|^2 = r.r
__global__ void k_pg_beta(const double* __restrict__ partials, int nb, int first, double* __restrict__ scal) {
  __shared__ double sh[kFinBlock / 64];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    a += partials[i];
    b += partials[kRedGrid + i];
  }
  const double rz = block_sum(a, sh), rr = block_sum(b, sh);
  if (threadIdx.x == 0) {
    scal[P_BETA] = first ? 0.0 : (scal[P_RHO] != 0.0 ? rz / scal[P_RHO] : 0.0);
    scal[P_RHO] = rz;
    scal[P_RR] = rr;
    if (first) scal[P_BB] = rr;
  }
}

// p = z + beta p
__global__ void k_pg_dir(int64_t n, const double* __restrict__ scal, const double* __restrict__ z,
                         double* __restrict__ p) {
  const double beta = scal[P_BETA];
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    p[c] = z[c] + beta * p[c];
}

// sum of x over the gauge cells and their count
__global__ void k_pg_gauge(int64_t n, const uint8_t* __restrict__ gauge, const double* __restrict__ x,
                           double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double s = 0.0, k = 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    if (gauge[c]) {
      s += x[c];
      k += 1.0;
    }
  const double a = block_sum(s, sh), b = block_sum(k, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a;
    partials[kRedGrid + blockIdx.x] = b;
  }
}

__global__ void k_pg_shift(int64_t n, const double* __restrict__ sums, double* __restrict__ x) {
  const double mean = sums[0] / sums[1];
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    x[c] -= mean;
}

// projection residual over the internal faces: sum w (p2 - p1 - j)^2 and sum w
__global__ void k_pg_mismatch(int64_t m, const int32_t* __restrict__ fc, const double* __restrict__ w,
                              const double* __restrict__ jmp, const double* __restrict__ x,
                              double* __restrict__ partials) {
  __shared__ double sh[kFinBlock / 64];
  double a = 0.0, b = 0.0;
  for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < m; f += (int64_t)gridDim.x * blockDim.x) {
    const int c2 = fc[2 * f + 1];
    if (c2 < 0) continue;
    const double wf = w[f];
    const double d = (x[c2] - x[fc[2 * f]]) - jmp[f];
    a += wf * d * d;
    b += wf;
  }
  const double ra = block_sum(a, sh), rb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = ra;
    partials[kRedGrid + blockIdx.x] = rb;
  }
}

}  // namespace
}  // namespace pyib

extern "C" {
using namespace pyib;

int pyib_pressure_graph(int64_t ncell, int64_t nface, const int64_t* face_cells, int32_t* ptr, int32_t* nbr,
                        int32_t* sface) {
  if (ncell < 2 || nface < 1 || ncell >= INT32_MAX || nface >= INT32_MAX)
    return fail_msg("pyib_pressure_graph: sizes outside [2, 2^31) cells / [1, 2^31) faces");
  for (int64_t c = 0; c <= ncell; ++c) ptr[c] = 0;
  int64_t nnz = 0;
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c1 = face_cells[2 * f], c2 = face_cells[2 * f + 1];
    if (c1 < 0 || c1 >= ncell || c2 < -1 || c2 >= ncell) return fail_msg("pyib_pressure_graph: invalid cell id");
    if (c2 < 0) continue;
    ++ptr[c1 + 1];
    ++ptr[c2 + 1];
    nnz += 2;
  }
  if (nnz >= INT32_MAX) return fail_msg("pyib_pressure_graph: too many internal faces for int32 offsets");
  for (int64_t c = 0; c < ncell; ++c) ptr[c + 1] += ptr[c];
  // each cell's internal faces in face order; sign + when the cell is the face's second owner
  int32_t* fill = new (std::nothrow) int32_t[ncell];
  if (!fill) return fail_msg("pyib_pressure_graph: out of host memory");
  for (int64_t c = 0; c < ncell; ++c) fill[c] = ptr[c];
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c1 = face_cells[2 * f], c2 = face_cells[2 * f + 1];
    if (c2 < 0) continue;
    int32_t k = fill[c1]++;
    nbr[k] = (int32_t)c2;
    sface[k] = (int32_t)(-f - 1);
    k = fill[c2]++;
    nbr[k] = (int32_t)c1;
    sface[k] = (int32_t)f;
  }
  delete[] fill;
  return 0;
}

int pyib_pressure_begin(const pyib_pgraph* g, void* stream) {
  auto st = S(stream);
  k_pg_face<<<grid_for(g->nface), kBlock, 0, st>>>(g->nface, g->fc, g->pos, g->grad, g->area, g->w, g->jmp);
  const int nb = red_grid_for(g->ncell);
  k_pg_begin<<<nb, kBlock, 0, st>>>(g->ncell, g->ptr, g->sface, g->w, g->jmp, g->dinv, g->x, g->r, g->z, g->p,
                                    g->partials);
  k_pg_beta<<<1, kFinBlock, 0, st>>>(g->partials, nb, 1, g->scal);
  return check_launch("pyib_pressure_begin");
}

int pyib_pressure_iterate(const pyib_pgraph* g, int32_t n_iter, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(g->ncell);
  for (int it = 0; it < n_iter; ++it) {
    k_pg_apply<<<nb, kBlock, 0, st>>>(g->ncell, g->ptr, g->nbr, g->sface, g->w, g->p, g->q, g->partials);
    k_pg_alpha<<<1, kFinBlock, 0, st>>>(g->partials, nb, g->scal);
    k_pg_update<<<nb, kBlock, 0, st>>>(g->ncell, g->scal, g->x, g->r, g->p, g->q, g->dinv, g->z, g->partials);
    k_pg_beta<<<1, kFinBlock, 0, st>>>(g->partials, nb, 0, g->scal);
    k_pg_dir<<<grid_for(g->ncell), kBlock, 0, st>>>(g->ncell, g->scal, g->z, g->p);
  }
  return check_launch("pyib_pressure_iterate");
}

int pyib_pressure_finish(const pyib_pgraph* g, double* out4, void* stream) {
  auto st = S(stream);
  const int nb = red_grid_for(g->ncell);
  k_pg_gauge<<<nb, kBlock, 0, st>>>(g->ncell, g->gauge, g->x, g->partials);
  k_finalize<<<1, kFinBlock, 0, st>>>(g->partials, nb, 2, out4);
  k_pg_shift<<<grid_for(g->ncell), kBlock, 0, st>>>(g->ncell, out4, g->x);
  const int nbf = red_grid_for(g->nface);
  k_pg_mismatch<<<nbf, kBlock, 0, st>>>(g->nface, g->fc, g->w, g->jmp, g->x, g->partials);
  k_finalize<<<1, kFinBlock, 0, st>>>(g->partials, nbf, 2, out4 + 2);
  return check_launch("pyib_pressure_finish");
}

}  // extern "C"
