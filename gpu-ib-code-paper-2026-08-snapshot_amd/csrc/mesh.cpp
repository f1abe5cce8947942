// mesh.cpp -- nested 2:1 Cartesian hex mesh in PyIB's global numbering (host setup).
//
// Produces the arrays of the reference's make_nested_mesh (src/pyib/make_euler_mesh.py:330-463)
// followed by its hanging-node finalisation (make_euler_mesh.py:787-950), in the same order:
//  * nodes are numbered coarsest level first, each level's new nodes in C order of its node
//    lattice; a node shared with a coarser level keeps the coarser id, and the coordinate of a
//    shared node is the one of the coarsest level holding it (the reference writes the levels
//    finest first, so the coarser write is the last);
//  * cells level by level, the live cells of each box (box minus the next finer box) in C order,
//    corners in the reference's corner order;
//  * faces in the lexicographic order of their sorted node ids, first owner = the smaller cell
//    id, node order = that owner's local face; edges in the lexicographic order of their
//    (smaller, larger) node ids;
//  * interface pairs (coarse face, four fine faces) per level pair and side in the reference's
//    emission order; fine interface faces take the coarse cell as second owner; the coarse
//    interface faces are removed; each coarse edge on the surface of a finer box is replaced in
//    its faces by its first half, and its second half is recorded in link_fe_insert.
//
// The reference builds faces by sorting the six faces of every cell and edges by a unique over
// the four edges of every face.  Here every level is a box lattice, so each face and edge is
// enumerated once from its lattice position, and the lexicographic numbering is a counting sort
// by the smallest node id followed by an insertion sort inside each (short) bucket.  The
// finalisation uses lattice positions instead of a geometric midpoint search.  The output is
// array-for-array the NumPy statement's, which the layout fingerprints of the reference's own
// D32/D40 runs pin.
#include <algorithm>
#include <array>
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

#ifdef PYIB_MESH_PROFILE
#include <chrono>
#include <cstdio>
// phase timings of the builder, compiled in only for profiling builds
static std::chrono::steady_clock::time_point mesh_t0 = std::chrono::steady_clock::now();
#define MESH_PHASE(name)                                                                              \
  std::fprintf(stderr, "[mesh] %-24s %8.3f s\n", name,                                              \
               std::chrono::duration<double>(std::chrono::steady_clock::now() - mesh_t0).count())
#else
#define MESH_PHASE(name) ((void)0)
#endif

namespace {

// corner offsets of a hex cell, in the reference's corner order (make_euler_mesh.py:210-213)
constexpr int CORNER[8][3] = {{1, 0, 1}, {1, 0, 0}, {0, 0, 1}, {0, 0, 0}, {1, 1, 1}, {1, 1, 0}, {0, 1, 1}, {0, 1, 0}};
// local faces as corner indices (make_euler_mesh.py:215-222): -y, +y, -z, +z, -x, +x
constexpr int HEXF[6][4] = {{3, 1, 0, 2}, {7, 5, 4, 6}, {3, 1, 5, 7}, {2, 0, 4, 6}, {3, 7, 6, 2}, {1, 5, 4, 0}};
// local face on the -/+ side of each axis
constexpr int LF_LO[3] = {4, 0, 2};
constexpr int LF_HI[3] = {5, 1, 3};

struct Lev {
  int64_t n[3];      // cells
  double lo[3], dx[3];
  int64_t hole[3];   // lower corner of the next finer box, in this level's cells (k >= 1)
  int64_t hn[3];     // extent of that box in this level's cells (0 on level 0)
  int64_t cell_off;  // first cell id of the level
  std::vector<int64_t> nid;   // node ids, (n0+1)(n1+1)(n2+1), C order
  std::vector<int64_t> cid;   // cell ids over the box, -1 in the hole
  std::vector<int64_t> fid[3];  // face ids (pre-removal) over the face lattice of each axis, -1 if none
  std::vector<int64_t> eid[3];  // edge ids (pre-collapse) over the edge lattice of each axis, -1 if none

  int64_t node(int64_t i, int64_t j, int64_t k) const { return nid[(i * (n[1] + 1) + j) * (n[2] + 1) + k]; }
  bool in_hole_cell(int64_t i, int64_t j, int64_t k) const {
    return hn[0] > 0 && i >= hole[0] && i < hole[0] + hn[0] && j >= hole[1] && j < hole[1] + hn[1] &&
           k >= hole[2] && k < hole[2] + hn[2];
  }
  bool live(int64_t i, int64_t j, int64_t k) const {
    return i >= 0 && j >= 0 && k >= 0 && i < n[0] && j < n[1] && k < n[2] && !in_hole_cell(i, j, k);
  }
  int64_t cell(int64_t i, int64_t j, int64_t k) const { return cid[(i * n[1] + j) * n[2] + k]; }
  // lattice extent of the faces normal to axis d / of the edges along axis d
  std::array<int64_t, 3> fshape(int d) const {
    return {n[0] + (d == 0), n[1] + (d == 1), n[2] + (d == 2)};
  }
  std::array<int64_t, 3> eshape(int d) const {
    return {n[0] + (d != 0), n[1] + (d != 1), n[2] + (d != 2)};
  }
};

struct Built {
  std::vector<double> nodes;
  std::vector<int64_t> cells, bfaces, bfam, f2n, f2c, e2n, f2e, bc_node, bc_edge, bc_face, pairs, insert;
  std::vector<int64_t> layer_cells, layer_nodes;
};

inline void sort4(int64_t* a) {
  for (int i = 1; i < 4; ++i) {
    int64_t v = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > v) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = v;
  }
}

// ranks of items by a key whose first component is a node id: counting sort by that id,
// then insertion sort of each bucket on the remaining components (lexicographic order)
template <int K>
std::vector<int64_t> lex_order(const std::vector<std::array<int64_t, K>>& key, int64_t nnode,
                               std::vector<int64_t>* bucket_start) {
  const int64_t n = static_cast<int64_t>(key.size());
  std::vector<int64_t> start(nnode + 1, 0);
  for (int64_t i = 0; i < n; ++i) ++start[key[i][0] + 1];
  for (int64_t v = 0; v < nnode; ++v) start[v + 1] += start[v];
  std::vector<int64_t> order(n);
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  for (int64_t i = 0; i < n; ++i) order[fill[key[i][0]]++] = i;
  for (int64_t v = 0; v < nnode; ++v) {
    const int64_t a = start[v], b = start[v + 1];
    for (int64_t i = a + 1; i < b; ++i) {
      const int64_t x = order[i];
      int64_t j = i - 1;
      while (j >= a && key[order[j]] > key[x]) {
        order[j + 1] = order[j];
        --j;
      }
      order[j + 1] = x;
    }
  }
  if (bucket_start) *bucket_start = std::move(start);
  return order;
}

int build(int nl, const int64_t* shape, const double* lo, const double* dx, const int64_t* hole, Built& out,
          std::string& err) {
  std::vector<Lev> L(nl);
  for (int k = 0; k < nl; ++k) {
    for (int d = 0; d < 3; ++d) {
      L[k].n[d] = shape[3 * k + d];
      L[k].lo[d] = lo[3 * k + d];
      L[k].dx[d] = dx[3 * k + d];
      L[k].hole[d] = k > 0 ? hole[3 * k + d] : 0;
      L[k].hn[d] = k > 0 ? shape[3 * (k - 1) + d] / 2 : 0;
      if (L[k].n[d] <= 0) {
        err = "pyib_mesh_build: empty level box";
        return 1;
      }
    }
    if (k > 0)
      for (int d = 0; d < 3; ++d)
        if (L[k].hole[d] < 1 || L[k].hole[d] + L[k].hn[d] > L[k].n[d] - 1 || shape[3 * (k - 1) + d] % 2) {
          err = "pyib_mesh_build: a finer box must lie strictly inside the next coarser box with even cell counts";
          return 1;
        }
  }
  // ---- node ids, coarsest level first (make_euler_mesh.py:343-401)
  int64_t nxt = 0;
  out.layer_nodes.assign(nl, 0);
  for (int k = nl - 1; k >= 0; --k) {
    Lev& v = L[k];
    const int64_t a = v.n[0] + 1, b = v.n[1] + 1, c = v.n[2] + 1;
    v.nid.assign(a * b * c, -1);
    if (k < nl - 1) {
      const Lev& u = L[k + 1];
      for (int64_t i = 0; i < a; i += 2)
        for (int64_t j = 0; j < b; j += 2)
          for (int64_t l = 0; l < c; l += 2)
            v.nid[(i * b + j) * c + l] = u.node(u.hole[0] + i / 2, u.hole[1] + j / 2, u.hole[2] + l / 2);
    }
    const int64_t first = nxt;
    for (int64_t i = 0; i < a; ++i)
      for (int64_t j = 0; j < b; ++j)
        for (int64_t l = 0; l < c; ++l) {
          int64_t& id = v.nid[(i * b + j) * c + l];
          if (id >= 0) continue;
          // nodes strictly inside the finer box belong to the finer level
          const bool inner = v.hn[0] > 0 && i > v.hole[0] && i < v.hole[0] + v.hn[0] && j > v.hole[1] &&
                             j < v.hole[1] + v.hn[1] && l > v.hole[2] && l < v.hole[2] + v.hn[2];
          if (!inner) id = nxt++;
        }
    out.layer_nodes[k] = nxt - first;
  }
  const int64_t nnode = nxt;
  out.nodes.assign(3 * nnode, 0.0);
  for (int k = 0; k < nl; ++k) {   // finest first: the coarsest level's coordinate is the last written
    const Lev& v = L[k];
    for (int64_t i = 0; i <= v.n[0]; ++i)
      for (int64_t j = 0; j <= v.n[1]; ++j)
        for (int64_t l = 0; l <= v.n[2]; ++l) {
          const int64_t id = v.node(i, j, l);
          if (id < 0) continue;
          double* p = &out.nodes[3 * id];
          p[0] = v.lo[0] + static_cast<double>(i) * v.dx[0];
          p[1] = v.lo[1] + static_cast<double>(j) * v.dx[1];
          p[2] = v.lo[2] + static_cast<double>(l) * v.dx[2];
        }
  }
  MESH_PHASE("nodes");
  // ---- cells (make_euler_mesh.py:403-415)
  int64_t ncell = 0;
  out.layer_cells.assign(nl, 0);
  for (int k = 0; k < nl; ++k) {
    Lev& v = L[k];
    v.cell_off = ncell;
    v.cid.assign(v.n[0] * v.n[1] * v.n[2], -1);
    for (int64_t i = 0; i < v.n[0]; ++i)
      for (int64_t j = 0; j < v.n[1]; ++j)
        for (int64_t l = 0; l < v.n[2]; ++l)
          if (!v.in_hole_cell(i, j, l)) v.cid[(i * v.n[1] + j) * v.n[2] + l] = ncell++;
    out.layer_cells[k] = ncell - v.cell_off;
  }
  out.cells.resize(8 * ncell);
  for (int k = 0; k < nl; ++k) {
    const Lev& v = L[k];
    for (int64_t i = 0; i < v.n[0]; ++i)
      for (int64_t j = 0; j < v.n[1]; ++j)
        for (int64_t l = 0; l < v.n[2]; ++l) {
          const int64_t c = v.cell(i, j, l);
          if (c < 0) continue;
          for (int p = 0; p < 8; ++p) {
            const int64_t id = v.node(i + CORNER[p][0], j + CORNER[p][1], l + CORNER[p][2]);
            if (id < 0) {
              err = "pyib_mesh_build: a live cell has an unnumbered corner";
              return 1;
            }
            out.cells[8 * c + p] = id;
          }
        }
  }
  MESH_PHASE("cells");
  // ---- faces: one per lattice face with at least one live cell (make_euler_mesh.py:534-603)
  struct FaceRec {
    int64_t c1, c2, lf;
  };
  std::vector<FaceRec> frec;
  std::vector<std::array<int64_t, 4>> fkey;
  std::vector<int64_t> fpos;   // (level, axis, lattice index) packed, to fill fid after sorting
  {
    int64_t cap = 0;   // every lattice face: an upper bound that avoids regrowing the arrays
    for (int k = 0; k < nl; ++k)
      for (int d = 0; d < 3; ++d) {
        const auto s = L[k].fshape(d);
        cap += s[0] * s[1] * s[2];
      }
    frec.reserve(cap);
    fkey.reserve(cap);
    fpos.reserve(cap);
  }
  for (int k = 0; k < nl; ++k) {
    Lev& v = L[k];
    for (int d = 0; d < 3; ++d) {
      const auto s = v.fshape(d);
      v.fid[d].assign(s[0] * s[1] * s[2], -1);
      for (int64_t i = 0; i < s[0]; ++i)
        for (int64_t j = 0; j < s[1]; ++j)
          for (int64_t l = 0; l < s[2]; ++l) {
            int64_t q[3] = {i, j, l};
            int64_t m[3] = {i, j, l};
            m[d] -= 1;
            const bool lo_live = v.live(m[0], m[1], m[2]), hi_live = v.live(q[0], q[1], q[2]);
            if (!lo_live && !hi_live) continue;
            FaceRec r;
            const int64_t* o;
            if (lo_live) {   // the lower cell has the smaller id within a level
              r.c1 = v.cell(m[0], m[1], m[2]);
              r.c2 = hi_live ? v.cell(q[0], q[1], q[2]) : -1;
              r.lf = LF_HI[d];
              o = m;
            } else {
              r.c1 = v.cell(q[0], q[1], q[2]);
              r.c2 = -1;
              r.lf = LF_LO[d];
              o = q;
            }
            std::array<int64_t, 4> kk;
            for (int t = 0; t < 4; ++t) {
              const int* cc = CORNER[HEXF[r.lf][t]];
              kk[t] = v.node(o[0] + cc[0], o[1] + cc[1], o[2] + cc[2]);
            }
            sort4(kk.data());
            frec.push_back(r);
            fkey.push_back(kk);
            fpos.push_back((((int64_t)k * 3 + d) << 40) | ((i * s[1] + j) * s[2] + l));
          }
    }
  }
  MESH_PHASE("faces enumerated");
  const int64_t nface0 = static_cast<int64_t>(frec.size());
  std::vector<int64_t> forder = lex_order<4>(fkey, nnode, nullptr);
  std::vector<int64_t>().swap(out.f2n);
  std::vector<int64_t> f2n(4 * nface0), f2c(2 * nface0);
  for (int64_t f = 0; f < nface0; ++f) {
    const int64_t src = forder[f];
    const FaceRec& r = frec[src];
    const int64_t* cn = &out.cells[8 * r.c1];
    for (int t = 0; t < 4; ++t) f2n[4 * f + t] = cn[HEXF[r.lf][t]];
    f2c[2 * f] = r.c1;
    f2c[2 * f + 1] = r.c2;
    const int64_t pk = fpos[src];
    const int lvl = static_cast<int>((pk >> 40) / 3), ax = static_cast<int>((pk >> 40) % 3);
    L[lvl].fid[ax][pk & ((int64_t(1) << 40) - 1)] = f;
  }
  MESH_PHASE("faces sorted");
  std::vector<FaceRec>().swap(frec);
  std::vector<std::array<int64_t, 4>>().swap(fkey);
  std::vector<int64_t>().swap(fpos);
  std::vector<int64_t>().swap(forder);
  // ---- edges: one per lattice edge of a live cell, in (smaller, larger) node id order
  std::vector<std::array<int64_t, 2>> ekey;
  std::vector<int64_t> epos;
  {
    int64_t cap = 0;
    for (int k = 0; k < nl; ++k)
      for (int d = 0; d < 3; ++d) {
        const auto s = L[k].eshape(d);
        cap += s[0] * s[1] * s[2];
      }
    ekey.reserve(cap);
    epos.reserve(cap);
  }
  for (int k = 0; k < nl; ++k) {
    Lev& v = L[k];
    for (int d = 0; d < 3; ++d) {
      const auto s = v.eshape(d);
      v.eid[d].assign(s[0] * s[1] * s[2], -1);
      const int a = (d + 1) % 3, b = (d + 2) % 3;
      for (int64_t i = 0; i < s[0]; ++i)
        for (int64_t j = 0; j < s[1]; ++j)
          for (int64_t l = 0; l < s[2]; ++l) {
            int64_t q[3] = {i, j, l};
            bool any = false;
            for (int sa = -1; sa <= 0 && !any; ++sa)
              for (int sb = -1; sb <= 0 && !any; ++sb) {
                int64_t m[3] = {i, j, l};
                m[a] += sa;
                m[b] += sb;
                any = v.live(m[0], m[1], m[2]);
              }
            if (!any) continue;
            int64_t e[3] = {i, j, l};
            e[d] += 1;
            const int64_t n1 = v.node(q[0], q[1], q[2]), n2 = v.node(e[0], e[1], e[2]);
            ekey.push_back({std::min(n1, n2), std::max(n1, n2)});
            epos.push_back((((int64_t)k * 3 + d) << 40) | ((i * s[1] + j) * s[2] + l));
          }
    }
  }
  MESH_PHASE("edges enumerated");
  const int64_t nedge0 = static_cast<int64_t>(ekey.size());
  std::vector<int64_t> ebucket;
  std::vector<int64_t> eorder = lex_order<2>(ekey, nnode, &ebucket);
  std::vector<int64_t> e2n(2 * nedge0);
  for (int64_t e = 0; e < nedge0; ++e) {
    const int64_t src = eorder[e];
    e2n[2 * e] = ekey[src][0];
    e2n[2 * e + 1] = ekey[src][1];
    const int64_t pk = epos[src];
    const int lvl = static_cast<int>((pk >> 40) / 3), ax = static_cast<int>((pk >> 40) % 3);
    L[lvl].eid[ax][pk & ((int64_t(1) << 40) - 1)] = e;
  }
  MESH_PHASE("edges sorted");
  std::vector<std::array<int64_t, 2>>().swap(ekey);
  std::vector<int64_t>().swap(epos);
  std::vector<int64_t>().swap(eorder);
  auto edge_of = [&](int64_t n1, int64_t n2) -> int64_t {
    const int64_t a = std::min(n1, n2), b = std::max(n1, n2);
    for (int64_t e = ebucket[a]; e < ebucket[a + 1]; ++e)
      if (e2n[2 * e + 1] == b) return e;
    return -1;
  };
  std::vector<int64_t> f2e(4 * nface0);
  for (int64_t f = 0; f < nface0; ++f)
    for (int t = 0; t < 4; ++t) {
      const int64_t e = edge_of(f2n[4 * f + t], f2n[4 * f + (t + 1) % 4]);
      if (e < 0) {
        err = "pyib_mesh_build: a face edge is missing";
        return 1;
      }
      f2e[4 * f + t] = e;
    }
  MESH_PHASE("f2e");
  // ---- outer boundary quads and families (make_euler_mesh.py:481-531)
  const Lev& g = L[nl - 1];
  const int64_t X = g.n[0], Y = g.n[1], Z = g.n[2];
  auto push_q = [&](int64_t a, int64_t b, int64_t c, int64_t d, int64_t fam) {
    out.bfaces.insert(out.bfaces.end(), {a, b, c, d});
    out.bfam.push_back(fam);
  };
  for (int64_t j = 0; j < Y; ++j)
    for (int64_t l = 0; l < Z; ++l) push_q(g.node(0, j, l + 1), g.node(0, j + 1, l + 1), g.node(0, j + 1, l), g.node(0, j, l), 1);
  for (int64_t j = 0; j < Y; ++j)
    for (int64_t l = 0; l < Z; ++l) push_q(g.node(X, j + 1, l + 1), g.node(X, j, l + 1), g.node(X, j, l), g.node(X, j + 1, l), 2);
  for (int64_t i = 0; i < X; ++i)
    for (int64_t j = 0; j < Y; ++j) push_q(g.node(i, j, Z), g.node(i + 1, j, Z), g.node(i + 1, j + 1, Z), g.node(i, j + 1, Z), 3);
  for (int64_t i = 0; i < X; ++i)
    for (int64_t j = 0; j < Y; ++j) push_q(g.node(i, j, 0), g.node(i, j + 1, 0), g.node(i + 1, j + 1, 0), g.node(i + 1, j, 0), 4);
  for (int64_t i = 0; i < X; ++i)
    for (int64_t l = 0; l < Z; ++l) push_q(g.node(i, Y, l + 1), g.node(i + 1, Y, l + 1), g.node(i + 1, Y, l), g.node(i, Y, l), 5);
  for (int64_t i = 0; i < X; ++i)
    for (int64_t l = 0; l < Z; ++l) push_q(g.node(i + 1, 0, l + 1), g.node(i, 0, l + 1), g.node(i, 0, l), g.node(i + 1, 0, l), 6);
  std::vector<int64_t> bc_face(nface0, 0);
  {
    // the face of each boundary quad, from its lattice position on the coarsest level
    auto face_at = [&](int d, int64_t i, int64_t j, int64_t l) {
      const auto s = g.fshape(d);
      return g.fid[d][(i * s[1] + j) * s[2] + l];
    };
    for (int64_t j = 0; j < Y; ++j)
      for (int64_t l = 0; l < Z; ++l) {
        bc_face[face_at(0, 0, j, l)] = 1;
        bc_face[face_at(0, X, j, l)] = 2;
      }
    for (int64_t i = 0; i < X; ++i)
      for (int64_t j = 0; j < Y; ++j) {
        bc_face[face_at(2, i, j, Z)] = 3;
        bc_face[face_at(2, i, j, 0)] = 4;
      }
    for (int64_t i = 0; i < X; ++i)
      for (int64_t l = 0; l < Z; ++l) {
        bc_face[face_at(1, i, Y, l)] = 5;
        bc_face[face_at(1, i, 0, l)] = 6;
      }
  }
  std::vector<int64_t> bc_node(nnode, 0), bc_edge(nedge0, 0);
  for (int64_t f = 0; f < nface0; ++f) {
    const int64_t fam = bc_face[f];
    if (!fam) continue;
    for (int t = 0; t < 4; ++t) {
      bc_node[f2n[4 * f + t]] = std::max(bc_node[f2n[4 * f + t]], fam);
      bc_edge[f2e[4 * f + t]] = std::max(bc_edge[f2e[4 * f + t]], fam);
    }
  }
  MESH_PHASE("boundary");
  // ---- interface pairs (make_euler_mesh.py:606-709), emission order of the reference
  struct Side {
    int axis;
    bool is_max;
    int fine_lf, coarse_lf;
  };
  constexpr Side SIDES[6] = {{0, false, 4, 5}, {0, true, 5, 4}, {1, false, 0, 1},
                             {1, true, 1, 0}, {2, false, 2, 3}, {2, true, 3, 2}};
  // face id of a cell's local face: the face lattice position on the cell's own level
  auto cell_face = [&](const Lev& v, int64_t i, int64_t j, int64_t l, int lf) {
    int d = 0;
    bool hi = false;
    for (int a = 0; a < 3; ++a) {
      if (LF_LO[a] == lf) d = a, hi = false;
      if (LF_HI[a] == lf) d = a, hi = true;
    }
    int64_t q[3] = {i, j, l};
    if (hi) q[d] += 1;
    const auto s = v.fshape(d);
    return v.fid[d][(q[0] * s[1] + q[1]) * s[2] + q[2]];
  };
  std::vector<int64_t>& pairs = out.pairs;
  for (int k = 0; k + 1 < nl; ++k) {
    const Lev& fv = L[k];
    const Lev& cv = L[k + 1];
    int64_t alo[3], ahi[3];
    for (int d = 0; d < 3; ++d) alo[d] = cv.hole[d], ahi[d] = cv.hole[d] + fv.n[d] / 2;
    for (const Side& sd : SIDES) {
      const int ax = sd.axis;
      int o0 = -1, o1 = -1;
      for (int a = 0; a < 3; ++a)
        if (a != ax) (o0 < 0 ? o0 : o1) = a;
      for (int64_t c0 = alo[o0]; c0 < ahi[o0]; ++c0)
        for (int64_t c1 = alo[o1]; c1 < ahi[o1]; ++c1) {
          int64_t ci[3];
          ci[ax] = sd.is_max ? ahi[ax] : alo[ax] - 1;
          ci[o0] = c0;
          ci[o1] = c1;
          if (!cv.live(ci[0], ci[1], ci[2])) {
            err = "pyib_mesh_build: interface coarse cell missing";
            return 1;
          }
          pairs.push_back(cell_face(cv, ci[0], ci[1], ci[2], sd.coarse_lf));
          for (int d0 = 0; d0 < 2; ++d0)
            for (int d1 = 0; d1 < 2; ++d1) {
              int64_t fi[3];
              fi[ax] = sd.is_max ? fv.n[ax] - 1 : 0;
              fi[o0] = (c0 - alo[o0]) * 2 + d0;
              fi[o1] = (c1 - alo[o1]) * 2 + d1;
              pairs.push_back(cell_face(fv, fi[0], fi[1], fi[2], sd.fine_lf));
            }
        }
    }
  }
  for (int64_t v : pairs)
    if (v < 0) {
      err = "pyib_mesh_build: interface face lookup failed";
      return 1;
    }
  const int64_t npair = static_cast<int64_t>(pairs.size()) / 5;
  MESH_PHASE("pairs");
  // ---- hanging-node finalisation (make_euler_mesh.py:787-950)
  std::vector<int64_t> fmap(nface0, 0), emap(nedge0, 0);
  std::vector<int64_t> sub1(nedge0, -1), sub2(nedge0, -1);
  if (npair) {
    std::vector<char> keep_face(nface0, 1);
    for (int64_t p = 0; p < npair; ++p) {
      const int64_t* row = &pairs[5 * p];
      for (int t = 1; t < 5; ++t) f2c[2 * row[t] + 1] = f2c[2 * row[0]];
      keep_face[row[0]] = 0;
    }
    // coarse edges on the surface of a finer box: the finer level has a node at their midpoint
    for (int k = 0; k + 1 < nl; ++k) {
      const Lev& fv = L[k];
      const Lev& cv = L[k + 1];
      int64_t alo[3], ahi[3];
      for (int d = 0; d < 3; ++d) alo[d] = cv.hole[d], ahi[d] = cv.hole[d] + fv.n[d] / 2;
      for (int d = 0; d < 3; ++d) {
        const auto s = cv.eshape(d);
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        for (int64_t i = alo[0]; i <= ahi[0]; ++i)
          for (int64_t j = alo[1]; j <= ahi[1]; ++j)
            for (int64_t l = alo[2]; l <= ahi[2]; ++l) {
              int64_t q[3] = {i, j, l};
              if (q[d] == ahi[d]) continue;
              const bool surface = q[a] == alo[a] || q[a] == ahi[a] || q[b] == alo[b] || q[b] == ahi[b];
              if (!surface) continue;
              const int64_t ce = cv.eid[d][(q[0] * s[1] + q[1]) * s[2] + q[2]];
              if (ce < 0) {
                err = "pyib_mesh_build: coarse interface edge missing";
                return 1;
              }
              int64_t fm[3];
              for (int t = 0; t < 3; ++t) fm[t] = 2 * (q[t] - alo[t]);
              fm[d] += 1;
              const int64_t mid = fv.node(fm[0], fm[1], fm[2]);
              const int64_t n1 = e2n[2 * ce], n2 = e2n[2 * ce + 1];
              const int64_t s1 = edge_of(n1, mid), s2 = edge_of(mid, n2);
              if (mid < 0 || s1 < 0 || s2 < 0) {
                err = "pyib_mesh_build: fine halves of a coarse interface edge missing";
                return 1;
              }
              sub1[ce] = s1;
              sub2[ce] = s2;
            }
      }
    }
    std::vector<int64_t> hit_face, hit_edge;
    for (int64_t f = 0; f < nface0; ++f) {
      if (!keep_face[f]) continue;
      for (int t = 0; t < 4; ++t) {
        const int64_t e = f2e[4 * f + t];
        if (sub1[e] < 0) continue;
        hit_face.push_back(f);
        hit_edge.push_back(sub2[e]);
        f2e[4 * f + t] = sub1[e];
      }
    }
    int64_t ne = 0, nf = 0;
    for (int64_t e = 0; e < nedge0; ++e) emap[e] = sub1[e] >= 0 ? -1 : ne++;
    for (int64_t f = 0; f < nface0; ++f) fmap[f] = keep_face[f] ? nf++ : -1;
    out.insert.resize(2 * hit_face.size());
    for (size_t i = 0; i < hit_face.size(); ++i) {
      out.insert[2 * i] = fmap[hit_face[i]];
      out.insert[2 * i + 1] = emap[hit_edge[i]];
      if (out.insert[2 * i] < 0 || out.insert[2 * i + 1] < 0) {
        err = "pyib_mesh_build: link_fe_insert references a removed face or edge";
        return 1;
      }
    }
  } else {
    for (int64_t e = 0; e < nedge0; ++e) emap[e] = e;
    for (int64_t f = 0; f < nface0; ++f) fmap[f] = f;
  }
  MESH_PHASE("finalisation");
  int64_t nface = 0, nedge = 0;
  for (int64_t f = 0; f < nface0; ++f) nface += fmap[f] >= 0;
  for (int64_t e = 0; e < nedge0; ++e) nedge += emap[e] >= 0;
  out.f2n.resize(4 * nface);
  out.f2c.resize(2 * nface);
  out.f2e.resize(4 * nface);
  out.bc_face.resize(nface);
  for (int64_t f = 0; f < nface0; ++f) {
    const int64_t g2 = fmap[f];
    if (g2 < 0) continue;
    for (int t = 0; t < 4; ++t) {
      out.f2n[4 * g2 + t] = f2n[4 * f + t];
      const int64_t e = emap[f2e[4 * f + t]];
      if (e < 0) {
        err = "pyib_mesh_build: a kept face still references a removed coarse edge";
        return 1;
      }
      out.f2e[4 * g2 + t] = e;
    }
    out.f2c[2 * g2] = f2c[2 * f];
    out.f2c[2 * g2 + 1] = f2c[2 * f + 1];
    out.bc_face[g2] = bc_face[f];
  }
  out.e2n.resize(2 * nedge);
  out.bc_edge.resize(nedge);
  for (int64_t e = 0; e < nedge0; ++e) {
    const int64_t g2 = emap[e];
    if (g2 < 0) continue;
    out.e2n[2 * g2] = e2n[2 * e];
    out.e2n[2 * g2 + 1] = e2n[2 * e + 1];
    out.bc_edge[g2] = bc_edge[e];
  }
  MESH_PHASE("output");
  out.bc_node = std::move(bc_node);
  return 0;
}

template <class T>
void put(T* dst, const std::vector<T>& src) {
  if (dst && !src.empty()) std::memcpy(dst, src.data(), sizeof(T) * src.size());
}

}  // namespace

extern "C" int pyib_mesh_build(int32_t nlevels, const int64_t* shape, const double* lo, const double* dx,
                               const int64_t* hole, void** handle, int64_t* counts) {
  if (!handle || !counts || !shape || !lo || !dx || !hole || nlevels < 1 || nlevels > 16)
    return pyib::fail_msg("pyib_mesh_build: bad arguments");
  *handle = nullptr;
  std::unique_ptr<Built> b(new (std::nothrow) Built());
  if (!b) return pyib::fail_msg("pyib_mesh_build: out of host memory");
  std::string err;
  int rc;
  try {
    rc = build(nlevels, shape, lo, dx, hole, *b, err);
  } catch (const std::bad_alloc&) {
    return pyib::fail_msg("pyib_mesh_build: out of host memory");
  }
  if (rc) return pyib::fail_msg(err.c_str());
  counts[0] = static_cast<int64_t>(b->nodes.size() / 3);
  counts[1] = static_cast<int64_t>(b->cells.size() / 8);
  counts[2] = static_cast<int64_t>(b->bfam.size());
  counts[3] = static_cast<int64_t>(b->f2c.size() / 2);
  counts[4] = static_cast<int64_t>(b->e2n.size() / 2);
  counts[5] = static_cast<int64_t>(b->pairs.size() / 5);
  counts[6] = static_cast<int64_t>(b->insert.size() / 2);
  *handle = b.release();
  return 0;
}

extern "C" int pyib_mesh_fetch(void* handle, double* nodes, int64_t* cells, int64_t* bfaces, int64_t* bfam,
                               int64_t* f2n, int64_t* f2c, int64_t* e2n, int64_t* f2e, int64_t* bc_node,
                               int64_t* bc_edge, int64_t* bc_face, int64_t* pairs, int64_t* insert,
                               int64_t* layer_cells, int64_t* layer_nodes) {
  if (!handle) return pyib::fail_msg("pyib_mesh_fetch: null handle");
  const Built& b = *static_cast<const Built*>(handle);
  put(nodes, b.nodes);
  put(cells, b.cells);
  put(bfaces, b.bfaces);
  put(bfam, b.bfam);
  put(f2n, b.f2n);
  put(f2c, b.f2c);
  put(e2n, b.e2n);
  put(f2e, b.f2e);
  put(bc_node, b.bc_node);
  put(bc_edge, b.bc_edge);
  put(bc_face, b.bc_face);
  put(pairs, b.pairs);
  put(insert, b.insert);
  put(layer_cells, b.layer_cells);
  put(layer_nodes, b.layer_nodes);
  return 0;
}

extern "C" int pyib_mesh_free(void* handle) {
  delete static_cast<Built*>(handle);
  return 0;
}
