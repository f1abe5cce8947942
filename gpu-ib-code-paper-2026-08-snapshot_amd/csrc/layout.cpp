// layout.cpp -- host-side structured layout of the elliptic operator (setup, float64).
//
// Renumbers a nested 2:1 mesh level by level into dense boxes: cells (nx, ny, nz) and three
// +axis-oriented edge blocks per level (x: (nx, ny+1, nz+1), y, z), classifies every row of
// the three passes of A = rot.integr.V.interp.curl (ib_core_np.py:168-175) as a
// constant-coefficient stencil row or a composite row, and assembles the composite rows
// (G1: cell <- edges, V: cell <- cells, G3: edge <- cells) from the mesh geometry.
//
// The composite rows are summed the way NumPy's add.reduceat sums a run of duplicate
// (row, column) entries: the first entry plus NumPy's pairwise sum of the rest (a plain
// loop below eight terms, eight interleaved partial sums up to 128, halving above).
// With -ffp-contract=off the coefficients are therefore bitwise those of the NumPy
// statement of the same construction.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/pyib_hip.h"

namespace pyib {
int fail_msg(const char* msg);
}

namespace {

constexpr int kMaxLevels = 8;

struct Lev {
  int64_t n[3];
  double h;
  double lo[3];
  int64_t hl[3], hh[3];   // hole [hl, hh) in this level's cells (empty on level 0)
  int64_t cell_off;
  int64_t e_off[3];
  int64_t eshape(int d, int a) const { return n[a] + (a == d ? 0 : 1); }
  int64_t esize(int d) const { return eshape(d, 0) * eshape(d, 1) * eshape(d, 2); }
  int64_t ncell() const { return n[0] * n[1] * n[2]; }
};

struct Comp {
  std::vector<int32_t> rows, ptr, col;
  std::vector<double> coef;   // ncomp values per entry
  int ncomp = 1;
};

struct Built {
  std::vector<int64_t> cell_map, edge_map, edge_inv;
  std::vector<int8_t> edge_sigma, sigma_int;
  std::vector<uint8_t> reg_cell, reg_edge, reg_g1;
  Comp g1, vis, g3;
  int64_t stats[3] = {0, 0, 0};   // regular V cells, regular G1 cells, regular edges
};

// NumPy's pairwise summation of n terms a[0], a[stride], ... (numpy/_core/src/umath/loops_utils.h)
double pairwise(const double* a, int64_t n, int64_t stride) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += a[i * stride];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j * stride];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[(i + j) * stride];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i * stride];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise(a, n2, stride) + pairwise(a + n2 * stride, n - n2, stride);
}

// |a - b| <= tol*scale, or a == b when tol == 0
inline bool close(double a, double b, double tol, double scale) {
  return tol > 0 ? std::fabs(a - b) <= tol * scale : a == b;
}

// A composite: the rows `sel` (ascending internal ids), each fed by entries (col, value[ncomp])
// listed in the order of the NumPy construction; duplicates of one (row, col) are summed.
struct Assembler {
  int ncomp;
  std::vector<int32_t> ecol;
  std::vector<double> eval;
  std::vector<int32_t> erow;                // position of the entry's row
  Assembler(int nc) : ncomp(nc) {}
  void add(int32_t rowpos, int32_t col, const double* v) {
    erow.push_back(rowpos);
    ecol.push_back(col);
    for (int k = 0; k < ncomp; ++k) eval.push_back(v[k]);
  }
  void finish(const std::vector<int64_t>& sel, Comp& out) {
    const int64_t nr = static_cast<int64_t>(sel.size());
    const int64_t ne = static_cast<int64_t>(erow.size());
    std::vector<int64_t> start(nr + 1, 0);
    for (int64_t e = 0; e < ne; ++e) start[erow[e] + 1]++;
    for (int64_t r = 0; r < nr; ++r) start[r + 1] += start[r];
    std::vector<int64_t> order(ne), fill(start.begin(), start.end() - 1);
    for (int64_t e = 0; e < ne; ++e) order[fill[erow[e]]++] = e;   // stable by row
    out.ncomp = ncomp;
    out.rows.resize(nr);
    out.ptr.assign(nr + 1, 0);
    out.col.clear();
    out.coef.clear();
    std::vector<double> tmp;
    for (int64_t r = 0; r < nr; ++r) {
      out.rows[r] = static_cast<int32_t>(sel[r]);
      int64_t* b = order.data() + start[r];
      int64_t* e = order.data() + start[r + 1];
      std::stable_sort(b, e, [&](int64_t x, int64_t y) { return ecol[x] < ecol[y]; });
      for (int64_t* g = b; g < e;) {
        int64_t* h = g;
        while (h < e && ecol[*h] == ecol[*g]) ++h;
        const int64_t m = h - g;
        out.col.push_back(ecol[*g]);
        for (int k = 0; k < ncomp; ++k) {
          double first = eval[*g * ncomp + k];
          if (m > 1) {
            tmp.resize(m - 1);
            for (int64_t q = 1; q < m; ++q) tmp[q - 1] = eval[g[q] * ncomp + k];
            first = first + pairwise(tmp.data(), m - 1, 1);
          }
          out.coef.push_back(first);
        }
        g = h;
      }
      out.ptr[r + 1] = static_cast<int32_t>(out.col.size());
    }
  }
};

struct In {
  int nlev;
  Lev lev[kMaxLevels];
  const int64_t* layer_cells;
  int64_t ncell_i, nedge_i;
  int64_t nnode;
  const double* nodes;
  int64_t nedge;
  const int64_t* e2n;
  const double* te;
  int64_t ncell, nface;
  const double *vol_ci, *nf, *rf1, *rf2, *aol, *af;
  const int64_t* f2c;
  const int32_t *curl_ptr, *curl_idx, *rot_ptr, *rot_idx, *cf_ptr, *cf_mid, *cf_face, *cf_other;
  int margin;
  double tol;
  int demote_neighbours;
};

inline bool in_hole(const Lev& L, int64_t i, int64_t j, int64_t k) {
  return i >= L.hl[0] && i < L.hh[0] && j >= L.hl[1] && j < L.hh[1] && k >= L.hl[2] && k < L.hh[2];
}

// structural regularity of a cell: `m` cells from the box boundary and from the hole
inline bool reg_cell_struct(const Lev& L, int64_t i, int64_t j, int64_t k, int64_t m) {
  if (i < m || i >= L.n[0] - m || j < m || j >= L.n[1] - m || k < m || k >= L.n[2] - m) return false;
  if (L.hh[0] > L.hl[0]) {
    const int64_t a0 = std::max<int64_t>(L.hl[0] - m, 0), a1 = std::max<int64_t>(L.hl[1] - m, 0),
                  a2 = std::max<int64_t>(L.hl[2] - m, 0);
    if (i >= a0 && i < L.hh[0] + m && j >= a1 && j < L.hh[1] + m && k >= a2 && k < L.hh[2] + m) return false;
  }
  return true;
}

// an edge of direction d at (i,j,k) of level L is regular when its four adjacent cells
// (offsets -1/0 on the two other axes) are all set in the level's cell mask `rc`
bool edge_cells_all(const Lev& L, const uint8_t* rc, int d, int64_t i, int64_t j, int64_t k) {
  const int o0 = d == 0 ? 1 : 0, o1 = d == 2 ? 1 : 2;
  const int64_t q[3] = {i, j, k};
  for (int s0 = -1; s0 <= 0; ++s0)
    for (int s1 = -1; s1 <= 0; ++s1) {
      int64_t c[3] = {q[0], q[1], q[2]};
      c[o0] += s0;
      c[o1] += s1;
      if (c[0] < 0 || c[0] >= L.n[0] || c[1] < 0 || c[1] >= L.n[1] || c[2] < 0 || c[2] >= L.n[2]) return false;
      if (!rc[(c[0] * L.n[1] + c[1]) * L.n[2] + c[2]]) return false;
    }
  return true;
}

int build(const In& in, Built& b, std::string& err) {
  const int N = in.nlev;
  const int64_t ncell = in.ncell, nface = in.nface, nedge = in.nedge;
  // ---- cells: reference order = C order of the live cells of each level
  b.cell_map.resize(ncell);
  {
    int64_t start = 0;
    for (int l = 0; l < N; ++l) {
      const Lev& L = in.lev[l];
      int64_t n = 0;
      for (int64_t i = 0; i < L.n[0]; ++i)
        for (int64_t j = 0; j < L.n[1]; ++j)
          for (int64_t k = 0; k < L.n[2]; ++k) {
            if (in_hole(L, i, j, k)) continue;
            if (start + n >= ncell) { err = "layer cell count does not match its box"; return 1; }
            b.cell_map[start + n] = L.cell_off + (i * L.n[1] + j) * L.n[2] + k;
            ++n;
          }
      if (n != in.layer_cells[l]) { err = "layer cell count does not match its box"; return 1; }
      start += n;
    }
    if (start != ncell) { err = "layer cell counts do not sum to the mesh's cells"; return 1; }
  }
  // ---- edges: level from the length, direction from the tangent, slot from the lower node
  b.edge_map.resize(nedge);
  b.edge_sigma.resize(nedge);
  const double h0 = in.lev[0].h;
  for (int64_t e = 0; e < nedge; ++e) {
    const double* t = in.te + 3 * e;
    int d = 0;
    double best = std::fabs(t[0]);
    for (int a = 1; a < 3; ++a)
      if (std::fabs(t[a]) > best) { best = std::fabs(t[a]); d = a; }
    const double length = best;
    const double lf = std::nearbyint(std::log2(length / h0));
    const int64_t l = static_cast<int64_t>(lf);
    const double want = h0 * std::exp2(lf);
    if (l < 0 || l >= N || !(std::fabs(length - want) <= 1e-8 + 1e-9 * std::fabs(want))) {
      err = "edge length is not a level spacing";
      return 1;
    }
    const int8_t sg = t[d] > 0 ? 1 : -1;
    b.edge_sigma[e] = sg;
    const double* lower = in.nodes + 3 * (sg > 0 ? in.e2n[2 * e] : in.e2n[2 * e + 1]);
    const Lev& L = in.lev[l];
    int64_t q[3];
    for (int a = 0; a < 3; ++a) {
      q[a] = static_cast<int64_t>(std::nearbyint((lower[a] - L.lo[a]) / L.h));
      if (q[a] < 0 || q[a] >= L.eshape(d, a)) { err = "edge outside its level box"; return 1; }
    }
    b.edge_map[e] = L.e_off[d] + (q[0] * L.eshape(d, 1) + q[1]) * L.eshape(d, 2) + q[2];
  }
  b.edge_inv.assign(in.nedge_i, -1);
  for (int64_t e = 0; e < nedge; ++e) {
    if (b.edge_inv[b.edge_map[e]] >= 0) { err = "two edges map to one slot"; return 1; }
    b.edge_inv[b.edge_map[e]] = e;
  }
  b.sigma_int.assign(in.nedge_i, 1);
  for (int64_t e = 0; e < nedge; ++e) b.sigma_int[b.edge_map[e]] = b.edge_sigma[e];
  // ---- structural classification (margin cells from the box and the hole)
  b.reg_cell.assign(in.ncell_i, 0);
  for (int l = 0; l < N; ++l) {
    const Lev& L = in.lev[l];
    for (int64_t i = 0; i < L.n[0]; ++i)
      for (int64_t j = 0; j < L.n[1]; ++j)
        for (int64_t k = 0; k < L.n[2]; ++k)
          b.reg_cell[L.cell_off + (i * L.n[1] + j) * L.n[2] + k] = reg_cell_struct(L, i, j, k, in.margin);
  }
  // ---- geometric regularity per pass
  std::vector<double> hc(ncell);
  for (int l = 0; l < N; ++l) {
    const Lev& L = in.lev[l];
    for (int64_t c = 0; c < ncell; ++c)
      if (b.cell_map[c] >= L.cell_off && b.cell_map[c] < L.cell_off + L.ncell()) hc[c] = L.h;
  }
  std::vector<uint8_t> ok_vol(ncell);
  for (int64_t c = 0; c < ncell; ++c) {
    const double ref = 1.0 / std::pow(hc[c], 3.0);
    ok_vol[c] = close(in.vol_ci[c], ref, in.tol, ref);
  }
  std::vector<uint8_t> rf1_ok(nface), rf2_ok(nface), links4(nface), v_face(nface), b_face(nface);
  for (int64_t f = 0; f < nface; ++f) {
    const double hf = hc[in.f2c[2 * f]];
    const double* n = in.nf + 3 * f;
    int dn = 0;
    double best = std::fabs(n[0]);
    for (int a = 1; a < 3; ++a)
      if (std::fabs(n[a]) > best) { best = std::fabs(n[a]); dn = a; }
    const double r1d = in.rf1[3 * f + dn];
    const bool sgn_nz = r1d > 0 || r1d < 0;
    links4[f] = (in.curl_ptr[f + 1] - in.curl_ptr[f]) == 4;
    const int64_t c2 = in.f2c[2 * f + 1];
    const bool m1 = c2 >= 0;
    const bool same_h = m1 && hc[std::max<int64_t>(c2, 0)] == hf;
    bool ok1 = sgn_nz, ok2 = same_h;
    for (int a = 0; a < 3; ++a) {
      const double half = 0.5 * hf * (a == dn ? 1.0 : 0.0);
      ok1 = ok1 && close(std::fabs(in.rf1[3 * f + a]), half, in.tol, hf);
      ok2 = ok2 && close(std::fabs(in.rf2[3 * f + a]), half, in.tol, hf);
    }
    rf1_ok[f] = ok1;
    rf2_ok[f] = ok2;
    v_face[f] = same_h && close(std::fabs(in.aol[f]), hf, in.tol, hf);
    b_face[f] = !m1 && close(in.aol[f], 2.0 * hf, in.tol, hf);   // boundary face, zero ghost, AoL = 2h
  }
  b.reg_g1.assign(in.ncell_i, 0);
  std::vector<uint8_t> real(in.ncell_i, 0), v_geo(in.ncell_i, 0), vb_geo(in.ncell_i, 0);
  for (int64_t c = 0; c < ncell; ++c) real[b.cell_map[c]] = 1;
  for (int64_t c = 0; c < ncell; ++c) {
    const int32_t s = in.cf_ptr[c], e = in.cf_ptr[c + 1], mid = in.cf_mid[c];
    bool g1 = ok_vol[c] && (e - s) == 6, v = g1, vb = g1;
    for (int32_t q = s; q < e; ++q) {
      const int32_t f = in.cf_face[q];
      const bool first = q < mid;
      if (!((first ? rf1_ok[f] : rf2_ok[f]) && links4[f])) g1 = false;
      if (!v_face[f]) v = false;
      if (!(in.cf_other[q] < ncell ? v_face[f] : b_face[f])) vb = false;
    }
    if (g1) b.reg_g1[b.cell_map[c]] = 1;
    if (v) v_geo[b.cell_map[c]] = 1;
    if (vb) vb_geo[b.cell_map[c]] = 1;
  }
  for (int64_t c = 0; c < in.ncell_i; ++c) b.reg_cell[c] = b.reg_cell[c] && v_geo[c] && real[c];
  // box-face V rows: cells of the coarsest level on the domain boundary, at least margin cells from
  // the hole, whose faces inside the box are same-level faces with AoL = h and whose other faces are
  // boundary faces with AoL = 2h (the stencil gives those 2 kA; reg_cell bit 1)
  std::vector<uint8_t> reg_vb(in.ncell_i, 0);
  {
    const Lev& L = in.lev[N - 1];
    const int64_t m = in.margin;
    for (int64_t i = 0; i < L.n[0]; ++i)
      for (int64_t j = 0; j < L.n[1]; ++j)
        for (int64_t k = 0; k < L.n[2]; ++k) {
          const bool onbox = i == 0 || i == L.n[0] - 1 || j == 0 || j == L.n[1] - 1 || k == 0 || k == L.n[2] - 1;
          if (!onbox) continue;
          if (L.hh[0] > L.hl[0] && i >= std::max<int64_t>(L.hl[0] - m, 0) && i < L.hh[0] + m &&
              j >= std::max<int64_t>(L.hl[1] - m, 0) && j < L.hh[1] + m &&
              k >= std::max<int64_t>(L.hl[2] - m, 0) && k < L.hh[2] + m)
            continue;
          const int64_t c = L.cell_off + (i * L.n[1] + j) * L.n[2] + k;
          reg_vb[c] = vb_geo[c] && real[c] && !b.reg_cell[c];
        }
  }
  for (int64_t c = 0; c < in.ncell_i; ++c) b.reg_cell[c] = b.reg_cell[c] || reg_vb[c];
  if (in.demote_neighbours) {
    std::vector<uint8_t> bad(in.ncell_i);
    for (int64_t c = 0; c < in.ncell_i; ++c) bad[c] = real[c] && !b.reg_cell[c];
    std::vector<uint8_t> nb(ncell, 0);
    for (int64_t c = 0; c < ncell; ++c)
      for (int32_t q = in.cf_ptr[c]; q < in.cf_ptr[c + 1]; ++q) {
        const int32_t o = in.cf_other[q];
        if (o < ncell && bad[b.cell_map[o]]) nb[c] = 1;
      }
    for (int64_t c = 0; c < ncell; ++c)
      if (nb[c]) b.reg_cell[b.cell_map[c]] = 0;
  }
  for (int64_t c = 0; c < in.ncell_i; ++c) b.reg_cell[c] = (uint8_t)((b.reg_cell[c] ? 1 : 0) | ((reg_vb[c] && b.reg_cell[c]) ? 2 : 0));
  // G3: four rot links on faces with closed-form rf1 and rf2, four real adjacent cells
  b.reg_edge.assign(in.nedge_i, 0);
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t s = in.rot_ptr[e], t = in.rot_ptr[e + 1];
    bool ok = (t - s) == 4;
    for (int32_t q = s; q < t && ok; ++q) {
      const int32_t v = in.rot_idx[q];
      const int32_t f = v >= 0 ? v : -v - 1;
      ok = rf1_ok[f] && rf2_ok[f] && links4[f];
    }
    if (ok) b.reg_edge[b.edge_map[e]] = 1;
  }
  for (int l = 0; l < N; ++l) {
    const Lev& L = in.lev[l];
    const uint8_t* rc = real.data() + L.cell_off;
    for (int d = 0; d < 3; ++d) {
      const int64_t s0 = L.eshape(d, 0), s1 = L.eshape(d, 1), s2 = L.eshape(d, 2);
      for (int64_t i = 0; i < s0; ++i)
        for (int64_t j = 0; j < s1; ++j)
          for (int64_t k = 0; k < s2; ++k) {
            const int64_t slot = L.e_off[d] + (i * s1 + j) * s2 + k;
            if (b.reg_edge[slot] && !edge_cells_all(L, rc, d, i, j, k)) b.reg_edge[slot] = 0;
          }
    }
  }
  for (int64_t e = 0; e < in.nedge_i; ++e)
    if (b.edge_inv[e] < 0) b.reg_edge[e] = 0;
  for (int64_t c = 0; c < in.ncell_i; ++c) {
    b.stats[0] += b.reg_cell[c] ? 1 : 0;
    b.stats[1] += b.reg_g1[c];
  }
  for (int64_t e = 0; e < in.nedge_i; ++e) b.stats[2] += b.reg_edge[e];
  // ---- composite rows
  std::vector<int64_t> cinv(in.ncell_i, -1);
  for (int64_t c = 0; c < ncell; ++c) cinv[b.cell_map[c]] = c;
  std::vector<int64_t> irr_g1, irr_v, irr_e;
  for (int64_t c = 0; c < in.ncell_i; ++c) {
    if (real[c] && !b.reg_g1[c]) irr_g1.push_back(c);
    if (real[c] && !b.reg_cell[c]) irr_v.push_back(c);
  }
  for (int64_t e = 0; e < in.nedge_i; ++e)
    if (b.edge_inv[e] >= 0 && !b.reg_edge[e]) irr_e.push_back(e);
  // G1 (cell <- edge): per cell-face entry, rfc = (rf1 | -rf2) * vol_ci, times the link sign
  // and the target edge's orientation, in cell-face then link order
  {
    Assembler as(3);
    for (size_t r = 0; r < irr_g1.size(); ++r) {
      const int64_t c = cinv[irr_g1[r]];
      for (int32_t q = in.cf_ptr[c]; q < in.cf_ptr[c + 1]; ++q) {
        const int32_t f = in.cf_face[q];
        const bool first = q < in.cf_mid[c];
        double rfc[3];
        for (int a = 0; a < 3; ++a)
          rfc[a] = (first ? in.rf1[3 * f + a] : -in.rf2[3 * f + a]) * in.vol_ci[c];
        for (int32_t k = in.curl_ptr[f]; k < in.curl_ptr[f + 1]; ++k) {
          const int32_t sv = in.curl_idx[k];
          const int64_t er = sv >= 0 ? sv : -static_cast<int64_t>(sv) - 1;
          const double s = (sv >= 0 ? 1.0 : -1.0) * static_cast<double>(b.edge_sigma[er]);
          const double v[3] = {rfc[0] * s, rfc[1] * s, rfc[2] * s};
          as.add(static_cast<int32_t>(r), static_cast<int32_t>(b.edge_map[er]), v);
        }
      }
    }
    as.finish(irr_g1, b.g1);
  }
  // V (cell <- cell): all off-diagonal entries vol_ci*AoL (inner faces), then all diagonal
  // entries -vol_ci*AoL, each group in cell-face order
  {
    Assembler as(1);
    for (int pass = 0; pass < 2; ++pass)
      for (size_t r = 0; r < irr_v.size(); ++r) {
        const int64_t c = cinv[irr_v[r]];
        for (int32_t q = in.cf_ptr[c]; q < in.cf_ptr[c + 1]; ++q) {
          const int32_t f = in.cf_face[q], o = in.cf_other[q];
          const double a = in.aol[f] * in.vol_ci[c];
          if (pass == 0) {
            if (o < ncell) as.add(static_cast<int32_t>(r), static_cast<int32_t>(b.cell_map[o]), &a);
          } else {
            const double na = -a;
            as.add(static_cast<int32_t>(r), static_cast<int32_t>(b.cell_map[c]), &na);
          }
        }
      }
    as.finish(irr_v, b.vis);
  }
  // G3 (edge <- cell): all first-owner entries rf1*s (edge then link order), then all
  // second-owner entries -rf2*s of the inner faces
  {
    Assembler as(3);
    for (int pass = 0; pass < 2; ++pass)
      for (size_t r = 0; r < irr_e.size(); ++r) {
        const int64_t er = b.edge_inv[irr_e[r]];
        for (int32_t k = in.rot_ptr[er]; k < in.rot_ptr[er + 1]; ++k) {
          const int32_t sv = in.rot_idx[k];
          const int64_t f = sv >= 0 ? sv : -static_cast<int64_t>(sv) - 1;
          const double s = (sv >= 0 ? 1.0 : -1.0) * static_cast<double>(b.edge_sigma[er]);
          const int64_t c1 = in.f2c[2 * f], c2 = in.f2c[2 * f + 1];
          if (pass == 0) {
            const double v[3] = {in.rf1[3 * f] * s, in.rf1[3 * f + 1] * s, in.rf1[3 * f + 2] * s};
            as.add(static_cast<int32_t>(r), static_cast<int32_t>(b.cell_map[c1]), v);
          } else if (c2 >= 0) {
            const double v[3] = {-in.rf2[3 * f] * s, -in.rf2[3 * f + 1] * s, -in.rf2[3 * f + 2] * s};
            as.add(static_cast<int32_t>(r), static_cast<int32_t>(b.cell_map[c2]), v);
          }
        }
      }
    as.finish(irr_e, b.g3);
  }
  return 0;
}

template <class T>
void put(T* dst, const std::vector<T>& src) {
  if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(T) * src.size());
}

}  // namespace

extern "C" int pyib_layout_build(int32_t nlevels, const int64_t* lev_i64, const double* lev_f64,
                                 const int64_t* layer_cells, int64_t ncell_i, int64_t nedge_i, int64_t nnode,
                                 const double* nodes, int64_t nedge, const int64_t* e2n, const double* te,
                                 int64_t ncell, int64_t nface, const double* vol_ci, const int64_t* f2c,
                                 const double* nf, const double* rf1, const double* rf2, const double* aol,
                                 const double* af, const int32_t* curl_ptr, const int32_t* curl_idx,
                                 const int32_t* rot_ptr, const int32_t* rot_idx, const int32_t* cf_ptr,
                                 const int32_t* cf_mid, const int32_t* cf_face, const int32_t* cf_other,
                                 int32_t margin, double tol, int32_t demote_neighbours, void** handle,
                                 int64_t* counts) {
  if (!handle || !counts || !lev_i64 || !lev_f64 || nlevels < 1 || nlevels > kMaxLevels || ncell <= 0 ||
      nedge <= 0 || nface <= 0 || margin < 0)
    return pyib::fail_msg("pyib_layout_build: bad arguments");
  if (ncell_i >= INT32_MAX || nedge_i >= INT32_MAX)
    return pyib::fail_msg("pyib_layout_build: mesh too large for int32 internal indices");
  *handle = nullptr;
  In in{};
  in.nlev = nlevels;
  for (int l = 0; l < nlevels; ++l) {
    const int64_t* q = lev_i64 + 13 * l;
    const double* g = lev_f64 + 4 * l;
    Lev& L = in.lev[l];
    for (int a = 0; a < 3; ++a) {
      L.n[a] = q[a];
      L.hl[a] = q[3 + a];
      L.hh[a] = q[6 + a];
      L.e_off[a] = q[10 + a];
      L.lo[a] = g[1 + a];
    }
    L.cell_off = q[9];
    L.h = g[0];
    if (L.n[0] <= 0 || L.n[1] <= 0 || L.n[2] <= 0 || L.cell_off < 0 || L.cell_off + L.ncell() > ncell_i)
      return pyib::fail_msg("pyib_layout_build: level box outside the internal cells");
    for (int d = 0; d < 3; ++d)
      if (L.e_off[d] < 0 || L.e_off[d] + L.esize(d) > nedge_i)
        return pyib::fail_msg("pyib_layout_build: level edge block outside the internal edges");
  }
  in.layer_cells = layer_cells;
  in.ncell_i = ncell_i;
  in.nedge_i = nedge_i;
  in.nnode = nnode;
  in.nodes = nodes;
  in.nedge = nedge;
  in.e2n = e2n;
  in.te = te;
  in.ncell = ncell;
  in.nface = nface;
  in.vol_ci = vol_ci;
  in.f2c = f2c;
  in.nf = nf;
  in.rf1 = rf1;
  in.rf2 = rf2;
  in.aol = aol;
  in.af = af;
  in.curl_ptr = curl_ptr;
  in.curl_idx = curl_idx;
  in.rot_ptr = rot_ptr;
  in.rot_idx = rot_idx;
  in.cf_ptr = cf_ptr;
  in.cf_mid = cf_mid;
  in.cf_face = cf_face;
  in.cf_other = cf_other;
  in.margin = margin;
  in.tol = tol;
  in.demote_neighbours = demote_neighbours;
  std::unique_ptr<Built> b(new (std::nothrow) Built());
  if (!b) return pyib::fail_msg("pyib_layout_build: out of host memory");
  std::string err;
  int rc;
  try {
    rc = build(in, *b, err);
  } catch (const std::bad_alloc&) {
    return pyib::fail_msg("pyib_layout_build: out of host memory");
  }
  if (rc) return pyib::fail_msg(("pyib_layout_build: " + err).c_str());
  const Comp* cs[3] = {&b->g1, &b->vis, &b->g3};
  for (int k = 0; k < 3; ++k) {
    if (cs[k]->col.size() >= static_cast<size_t>(INT32_MAX))
      return pyib::fail_msg("pyib_layout_build: composite too large for int32 offsets");
    counts[2 * k] = static_cast<int64_t>(cs[k]->rows.size());
    counts[2 * k + 1] = static_cast<int64_t>(cs[k]->col.size());
  }
  counts[6] = b->stats[0];
  counts[7] = b->stats[1];
  counts[8] = b->stats[2];
  *handle = b.release();
  return 0;
}

extern "C" int pyib_layout_fetch(void* handle, int64_t* cell_map, int64_t* edge_map, int8_t* edge_sigma,
                                 int64_t* edge_inv, int8_t* sigma_int, uint8_t* reg_cell, uint8_t* reg_edge,
                                 uint8_t* reg_g1, int32_t* g1_rows, int32_t* g1_ptr, int32_t* g1_col,
                                 double* g1_coef, int32_t* v_rows, int32_t* v_ptr, int32_t* v_col, double* v_coef,
                                 int32_t* g3_rows, int32_t* g3_ptr, int32_t* g3_col, double* g3_coef) {
  if (!handle) return pyib::fail_msg("pyib_layout_fetch: null handle");
  const Built& b = *static_cast<const Built*>(handle);
  put(cell_map, b.cell_map);
  put(edge_map, b.edge_map);
  put(edge_sigma, b.edge_sigma);
  put(edge_inv, b.edge_inv);
  put(sigma_int, b.sigma_int);
  put(reg_cell, b.reg_cell);
  put(reg_edge, b.reg_edge);
  put(reg_g1, b.reg_g1);
  put(g1_rows, b.g1.rows);
  put(g1_ptr, b.g1.ptr);
  put(g1_col, b.g1.col);
  put(g1_coef, b.g1.coef);
  put(v_rows, b.vis.rows);
  put(v_ptr, b.vis.ptr);
  put(v_col, b.vis.col);
  put(v_coef, b.vis.coef);
  put(g3_rows, b.g3.rows);
  put(g3_ptr, b.g3.ptr);
  put(g3_col, b.g3.col);
  put(g3_coef, b.g3.coef);
  return 0;
}

extern "C" int pyib_layout_free(void* handle) {
  delete static_cast<Built*>(handle);
  return 0;
}
