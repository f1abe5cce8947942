// geometry.cpp -- host-side mesh geometry of the stream-function operators (setup, float64).
//
// The quantities of the reference's IBCore.__init__ (ib_core_np.py:24-126): the flat
// face/edge link list (4 links per face + link_fe_insert), edge midpoints and tangents,
// face and cell centres, the per-link circulation sign, face area vectors, the 2:1
// "Modify_geom" centre corrections (ib_core_np.py:74-102), rf/rf1/rf2, vol_ci and AoL.
//
// Every sum runs in the order of the reference's np.add.at / np.bincount over the link
// list (and over [c1 owners, then c2 neighbours] for cells), every 3-vector dot product is
// ((a0 b0 + a1 b1) + a2 b2), and the library is built with -ffp-contract=off, so each value
// is bitwise the one the NumPy statement of the same algorithm produces.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pyib_hip.h"

namespace pyib {
int fail_msg(const char* msg);   // sets the message pyib_last_error() returns
}

namespace {

inline double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// cell average of a face vector: sum over the faces owning the cell (c1, in face order),
// then over the faces neighbouring it (c2, in face order), divided by the count
void cells_avg(int64_t nface, int64_t ncell, const int64_t* f2c, const double* fv, const double* ccnt,
               double* out) {
  std::memset(out, 0, sizeof(double) * 3 * ncell);
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c = f2c[2 * f];
    for (int d = 0; d < 3; ++d) out[3 * c + d] += fv[3 * f + d];
  }
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c = f2c[2 * f + 1];
    if (c < 0) continue;
    for (int d = 0; d < 3; ++d) out[3 * c + d] += fv[3 * f + d];
  }
  for (int64_t c = 0; c < ncell; ++c)
    for (int d = 0; d < 3; ++d) out[3 * c + d] /= ccnt[c];
}

}  // namespace

extern "C" int pyib_geometry(int64_t nnode, const double* nodes, int64_t nedge, const int64_t* e2n, int64_t nface,
                             const int64_t* f2c, const int64_t* f2e, int64_t nins, const int64_t* ins,
                             int64_t ncell, int64_t* lf_face, int64_t* lf_edge, double* epos, double* te,
                             double* fpos, double* cpos, double* sgn, double* nf, double* af, double* rf1,
                             double* rf2, double* rf, double* vol_ci, double* aol) {
  if (nnode <= 0 || nedge <= 0 || nface <= 0 || ncell <= 0 || nins < 0)
    return pyib::fail_msg("pyib_geometry: empty mesh");
  const int64_t nlink = 4 * nface + nins;
  for (int64_t f = 0; f < nface; ++f)
    for (int k = 0; k < 4; ++k) {
      lf_face[4 * f + k] = f;
      lf_edge[4 * f + k] = f2e[4 * f + k];
    }
  for (int64_t i = 0; i < nins; ++i) {
    lf_face[4 * nface + i] = ins[2 * i];
    lf_edge[4 * nface + i] = ins[2 * i + 1];
  }
  for (int64_t e = 0; e < nedge; ++e) {
    const double* a = nodes + 3 * e2n[2 * e];
    const double* b = nodes + 3 * e2n[2 * e + 1];
    for (int d = 0; d < 3; ++d) {
      epos[3 * e + d] = 0.5 * (a[d] + b[d]);
      te[3 * e + d] = b[d] - a[d];
    }
  }
  // face centres: mean of the link edge midpoints (link order)
  std::vector<double> fcnt(nface, 0.0);
  std::memset(fpos, 0, sizeof(double) * 3 * nface);
  for (int64_t l = 0; l < nlink; ++l) {
    const int64_t f = lf_face[l], e = lf_edge[l];
    fcnt[f] += 1.0;
    for (int d = 0; d < 3; ++d) fpos[3 * f + d] += epos[3 * e + d];
  }
  for (int64_t f = 0; f < nface; ++f)
    for (int d = 0; d < 3; ++d) fpos[3 * f + d] /= fcnt[f];
  std::vector<double> ccnt(ncell, 0.0);
  for (int64_t f = 0; f < nface; ++f) ccnt[f2c[2 * f]] += 1.0;
  for (int64_t f = 0; f < nface; ++f)
    if (f2c[2 * f + 1] >= 0) ccnt[f2c[2 * f + 1]] += 1.0;
  cells_avg(nface, ncell, f2c, fpos, ccnt.data(), cpos);
  // centre-to-centre (or centre-to-face on the boundary) vector, for the circulation sign
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
    for (int d = 0; d < 3; ++d)
      rf[3 * f + d] = c2 >= 0 ? cpos[3 * c2 + d] - cpos[3 * c1 + d] : fpos[3 * f + d] - cpos[3 * c1 + d];
  }
  // link area vector (edge-to-face-centre x tangent), its sign against rf, face area vectors
  std::memset(nf, 0, sizeof(double) * 3 * nface);
  for (int64_t l = 0; l < nlink; ++l) {
    const int64_t f = lf_face[l], e = lf_edge[l];
    const double a0 = epos[3 * e] - fpos[3 * f], a1 = epos[3 * e + 1] - fpos[3 * f + 1],
                 a2 = epos[3 * e + 2] - fpos[3 * f + 2];
    const double* t = te + 3 * e;
    const double av[3] = {a1 * t[2] - a2 * t[1], a2 * t[0] - a0 * t[2], a0 * t[1] - a1 * t[0]};
    const double s = dot3(av, rf + 3 * f);
    const double sg = s > 0.0 ? 1.0 : (s < 0.0 ? -1.0 : 1.0);
    sgn[l] = sg;
    for (int d = 0; d < 3; ++d) nf[3 * f + d] += av[d] * sg;
  }
  for (int64_t i = 0; i < 3 * nface; ++i) nf[i] = nf[i] * 0.5;
  for (int64_t f = 0; f < nface; ++f) af[f] = std::sqrt(dot3(nf + 3 * f, nf + 3 * f));
  // 2:1 interface face centres: shift by rfe/6 on the links with |rfe|^2 = 0.36 Af (tested on
  // the unshifted centres, applied afterwards in link order)
  {
    std::vector<int64_t> hit;
    std::vector<double> inc;
    for (int64_t l = 0; l < nlink; ++l) {
      const int64_t f = lf_face[l], e = lf_edge[l];
      const double r[3] = {epos[3 * e] - fpos[3 * f], epos[3 * e + 1] - fpos[3 * f + 1],
                           epos[3 * e + 2] - fpos[3 * f + 2]};
      if (std::fabs(dot3(r, r) - 0.36 * af[f]) < 1e-6 * af[f]) {
        hit.push_back(f);
        for (int d = 0; d < 3; ++d) inc.push_back(r[d] / 6.0);
      }
    }
    for (size_t i = 0; i < hit.size(); ++i)
      for (int d = 0; d < 3; ++d) fpos[3 * hit[i] + d] += inc[3 * i + d];
  }
  cells_avg(nface, ncell, f2c, fpos, ccnt.data(), cpos);
  // 2:1 interface cell centres: shift by rf/4 where |rf|^2 = 4/9 Af (owners first, then
  // neighbours, both tested on the unshifted centres)
  {
    const double k49 = 4.0 / 9.0;
    std::vector<int64_t> h1c, h2c;
    std::vector<double> h1v, h2v;
    for (int64_t f = 0; f < nface; ++f) {
      const int64_t c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
      const double tol = 1e-6 * af[f], a49 = k49 * af[f];
      const double r1[3] = {fpos[3 * f] - cpos[3 * c1], fpos[3 * f + 1] - cpos[3 * c1 + 1],
                            fpos[3 * f + 2] - cpos[3 * c1 + 2]};
      const bool h1 = std::fabs(dot3(r1, r1) - a49) < tol;
      if (h1) {
        h1c.push_back(c1);
        for (int d = 0; d < 3; ++d) h1v.push_back(r1[d] / 4.0);
      } else if (c2 >= 0) {
        const double r2[3] = {fpos[3 * f] - cpos[3 * c2], fpos[3 * f + 1] - cpos[3 * c2 + 1],
                              fpos[3 * f + 2] - cpos[3 * c2 + 2]};
        if (std::fabs(dot3(r2, r2) - a49) < tol) {
          h2c.push_back(c2);
          for (int d = 0; d < 3; ++d) h2v.push_back(r2[d] / 4.0);
        }
      }
    }
    for (size_t i = 0; i < h1c.size(); ++i)
      for (int d = 0; d < 3; ++d) cpos[3 * h1c[i] + d] += h1v[3 * i + d];
    for (size_t i = 0; i < h2c.size(); ++i)
      for (int d = 0; d < 3; ++d) cpos[3 * h2c[i] + d] += h2v[3 * i + d];
  }
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
    for (int d = 0; d < 3; ++d) {
      rf1[3 * f + d] = fpos[3 * f + d] - cpos[3 * c1 + d];
      rf2[3 * f + d] = c2 >= 0 ? fpos[3 * f + d] - cpos[3 * c2 + d] : 0.0;
      rf[3 * f + d] = rf1[3 * f + d] - rf2[3 * f + d];
    }
  }
  // vol_ci = 3 / sum of nf.rf over the cell's faces (owners, then neighbours)
  std::vector<double> vraw(ncell, 0.0);
  for (int64_t f = 0; f < nface; ++f) vraw[f2c[2 * f]] += dot3(nf + 3 * f, rf1 + 3 * f);
  for (int64_t f = 0; f < nface; ++f)
    if (f2c[2 * f + 1] >= 0) vraw[f2c[2 * f + 1]] += -dot3(nf + 3 * f, rf2 + 3 * f);
  for (int64_t c = 0; c < ncell; ++c) vol_ci[c] = std::fabs(vraw[c]) > 1e-14 ? 3.0 / vraw[c] : 0.0;
  for (int64_t f = 0; f < nface; ++f) aol[f] = dot3(rf + 3 * f, nf + 3 * f) / dot3(rf + 3 * f, rf + 3 * f);
  return 0;
}

// Gather tables of the incidence operators (the CSR the gather kernels read), by stable
// counting sorts: the curl table lists each face's links in link order, the rot table each
// edge's links in link order, and the cell-face table each cell's faces as first owner, then
// as second owner, each in face order (the order of the reference's bincount(c1) - bincount(c2)).
// Signs are folded into the index: v >= 0 -> +x[v], v < 0 -> -x[-v-1].  Boundary face k (k-th
// in face order) has the ghost index ncell + k as its opposite owner.
extern "C" int pyib_tables(int64_t nface, int64_t nedge, int64_t ncell, int64_t nlink, const int64_t* lf_face,
                           const int64_t* lf_edge, const double* sgn, const int64_t* f2c, int32_t* curl_ptr,
                           int32_t* curl_idx, int32_t* rot_ptr, int32_t* rot_idx, int32_t* cf_ptr, int32_t* cf_mid,
                           int32_t* cf_face, int32_t* cf_other) {
  if (nface >= INT32_MAX || nedge >= INT32_MAX || ncell >= INT32_MAX || nlink >= INT32_MAX ||
      2 * nface >= INT32_MAX)
    return pyib::fail_msg("pyib_tables: a size does not fit the int32 tables");
  auto enc = [](int64_t v, double s) { return (int32_t)(s > 0.0 ? v : -v - 1); };
  // curl: bucket the links by face
  std::vector<int64_t> pos(nface + 1, 0);
  for (int64_t l = 0; l < nlink; ++l) ++pos[lf_face[l] + 1];
  for (int64_t f = 0; f < nface; ++f) pos[f + 1] += pos[f];
  for (int64_t f = 0; f <= nface; ++f) curl_ptr[f] = (int32_t)pos[f];
  for (int64_t l = 0; l < nlink; ++l) curl_idx[pos[lf_face[l]]++] = enc(lf_edge[l], sgn[l]);
  // rot: bucket the links by edge
  pos.assign(nedge + 1, 0);
  for (int64_t l = 0; l < nlink; ++l) ++pos[lf_edge[l] + 1];
  for (int64_t e = 0; e < nedge; ++e) pos[e + 1] += pos[e];
  for (int64_t e = 0; e <= nedge; ++e) rot_ptr[e] = (int32_t)pos[e];
  for (int64_t l = 0; l < nlink; ++l) rot_idx[pos[lf_edge[l]]++] = enc(lf_face[l], sgn[l]);
  // cell <- face: first-owner faces, then second-owner faces, each in face order
  std::vector<int64_t> n0(ncell, 0), n1(ncell, 0);
  for (int64_t f = 0; f < nface; ++f) {
    ++n0[f2c[2 * f]];
    if (f2c[2 * f + 1] >= 0) ++n1[f2c[2 * f + 1]];
  }
  std::vector<int64_t> p0(ncell), p1(ncell);
  int64_t acc = 0;
  for (int64_t c = 0; c < ncell; ++c) {
    cf_ptr[c] = (int32_t)acc;
    cf_mid[c] = (int32_t)(acc + n0[c]);
    p0[c] = acc;
    p1[c] = acc + n0[c];
    acc += n0[c] + n1[c];
  }
  cf_ptr[ncell] = (int32_t)acc;
  int64_t ghost = ncell;
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c1 = f2c[2 * f], c2 = f2c[2 * f + 1];
    const int64_t i = p0[c1]++;
    cf_face[i] = (int32_t)f;
    cf_other[i] = (int32_t)(c2 >= 0 ? c2 : ghost++);
  }
  for (int64_t f = 0; f < nface; ++f) {
    const int64_t c2 = f2c[2 * f + 1];
    if (c2 < 0) continue;
    const int64_t i = p1[c2]++;
    cf_face[i] = (int32_t)f;
    cf_other[i] = (int32_t)f2c[2 * f];
  }
  return 0;
}
