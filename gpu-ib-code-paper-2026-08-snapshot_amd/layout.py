"""Structured per-level layout of the elliptic operator (host setup, NumPy).

Each 2:1 level L is a box of (nx, ny, nz) cells of spacing h_L.  The CG
unknowns are renumbered level by level into three dense edge blocks
(x-edges (nx, ny+1, nz+1), y-edges (nx+1, ny, nz+1), z-edges (nx+1, ny+1, nz))
and cell vectors into (nx, ny, nz) blocks, all with the +axis orientation
(``sigma`` = +-1 maps PyIB's edge orientation, node order of E2N, onto it).
Slots of the box that are not mesh elements (the hole under a finer level,
coarse edges collapsed onto fine halves) stay zero.

Away from level boundaries the operator A = rot.integr.V.interp.curl
(ib_core_np.py:168-175) is a constant-coefficient stencil: the "regular"
rows.  They are cells at least ``margin`` cells from the box boundary and
from the hole, and edges whose four adjacent cells are regular.  A regular
row is kept only if the mesh geometry around it equals the closed form
(rf1 = +-h/2 e_d, AoL = h, vol_ci = 1/h^3, Af = h^2, 4 unit-sign links) to
the tolerance of the arithmetic precision; otherwise it is demoted.  Every
other row ("irregular": level interfaces, the outer boundary) is an explicit
composite row built here from the reference geometry:

* G1 (cell <- edges): vol_ci * (sum_c1 rf1 U - sum_c2 rf2 U), U = curl p
* V  (cell <- cells): x - ...: coefficients vol_ci*AoL per face, diagonal
  -vol_ci*sum AoL (zero ghost), applied as y = rou x - (hdt*visc) V x
* G3 (edge <- cells): rot(integr(.)) with the two owners' rf1/-rf2
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MAX_LEVELS = 8


@dataclass
class Level:
    nx: int
    ny: int
    nz: int
    h: float
    lo: tuple
    hole_lo: tuple
    hole_hi: tuple
    cell_off: int
    e_off: tuple
    ncell: int = 0
    nedge: int = 0


@dataclass
class Composite:
    rows: np.ndarray            # internal row ids
    ptr: np.ndarray
    col: np.ndarray
    coef: np.ndarray            # (nnz,) or (nnz,3)


@dataclass
class StructuredLayout:
    levels: list
    ncell: int                  # internal (padded) cells
    nedge: int                  # internal (padded) edges
    cell_map: np.ndarray        # reference cell -> internal
    edge_map: np.ndarray        # reference edge -> internal
    edge_sigma: np.ndarray      # +-1
    edge_inv: np.ndarray        # internal -> reference edge (-1: empty slot)
    sigma_int: np.ndarray       # +-1 on internal slots (1 on empty)
    reg_cell: np.ndarray        # uint8 over internal cells: bit 0 V stencil row, bit 1 box-face V row
    reg_edge: np.ndarray        # uint8 over internal edges: G3 stencil row
    reg_g1: np.ndarray = None   # uint8 over internal cells: G1 stencil row
    g1: Composite = None
    vis: Composite = None
    g3: Composite = None
    stats: dict = field(default_factory=dict)


def _edge_shape(lv, d):
    n = [lv.nx, lv.ny, lv.nz]
    return tuple(n[a] + (0 if a == d else 1) for a in range(3))


def _cells_of_level(lv):
    m = np.ones((lv.nx, lv.ny, lv.nz), bool)
    a, b = lv.hole_lo, lv.hole_hi
    m[a[0]:b[0], a[1]:b[1], a[2]:b[2]] = False
    return m


def _levels(mesh):
    """Level boxes and the internal numbering's offsets (cells, then the x/y/z edge blocks of
    each level in turn)."""
    N = mesh.n_layers
    if N > MAX_LEVELS:
        raise ValueError(f"at most {MAX_LEVELS} levels")
    levels = []
    coff, eoff = 0, 0
    for L in range(N):
        nx, ny, nz = (int(v) for v in mesh.layer_shape[L])
        if L > 0:
            hl = tuple(int(v) for v in mesh.layer_hole_lo[L])
            hs = tuple(int(v) // 2 for v in mesh.layer_shape[L - 1])
            hh = tuple(hl[a] + hs[a] for a in range(3))
        else:
            hl, hh = (0, 0, 0), (0, 0, 0)
        lv = Level(nx, ny, nz, float(mesh.layer_spacing[L]), tuple(float(v) for v in mesh.layer_min[L]), hl, hh,
                   coff, (0, 0, 0))
        lv.ncell = nx * ny * nz
        offs = []
        for d in range(3):
            offs.append(eoff)
            eoff += int(np.prod(_edge_shape(lv, d)))
        lv.e_off = tuple(offs)
        coff += lv.ncell
        levels.append(lv)
    if max(coff, eoff) >= 2 ** 31 - 1:
        raise ValueError("mesh too large for int32 internal indices")
    return levels, coff, eoff


def build_layout(mesh, geo, tables, margin=1, tol=1e-13, demote_neighbours=False):
    """Internal layout + classification + composite rows, built in C++ (``pyib_layout_build`` in
    ``csrc/layout.cpp``); array for array equal to ``build_layout_numpy`` (the same arguments)."""
    import ctypes as C
    from . import _native as N
    levels, ncell_i, nedge_i = _levels(mesh)
    li = np.array([[lv.nx, lv.ny, lv.nz, *lv.hole_lo, *lv.hole_hi, lv.cell_off, *lv.e_off] for lv in levels],
                  np.int64)
    lf = np.array([[lv.h, *lv.lo] for lv in levels], np.float64)
    c64 = lambda a: np.ascontiguousarray(a, np.float64)   # noqa: E731
    i64 = lambda a: np.ascontiguousarray(a, np.int64)     # noqa: E731
    i32 = lambda a: np.ascontiguousarray(a, np.int32)     # noqa: E731
    keep = [li, lf, i64(mesh.layer_cell_counts), c64(mesh.nodes), i64(mesh.edges_to_nodes), c64(geo.te),
            c64(geo.vol_ci), i64(geo.F2C), c64(geo.nf), c64(geo.rf1), c64(geo.rf2), c64(geo.AoL), c64(geo.Af),
            i32(tables.curl_ptr), i32(tables.curl_idx), i32(tables.rot_ptr), i32(tables.rot_idx),
            i32(tables.cf_ptr), i32(tables.cf_mid), i32(tables.cf_face), i32(tables.cf_other)]
    p = lambda a: a.ctypes.data_as(C.c_void_p)            # noqa: E731
    (li, lf, lcc, nodes, e2n, te, vol, f2c, nf, rf1, rf2, aol, af, cp, ci, rp, ri, fp, fm, ff, fo) = keep
    handle = C.c_void_p()
    counts = np.zeros(9, np.int64)
    N.call("pyib_layout_build", C.c_int32(len(levels)), p(li), p(lf), p(lcc), C.c_int64(ncell_i),
           C.c_int64(nedge_i), C.c_int64(nodes.shape[0]), p(nodes), C.c_int64(e2n.shape[0]), p(e2n), p(te),
           C.c_int64(geo.ncell), C.c_int64(geo.nface), p(vol), p(f2c), p(nf), p(rf1), p(rf2), p(aol), p(af),
           p(cp), p(ci), p(rp), p(ri), p(fp), p(fm), p(ff), p(fo), C.c_int32(margin), C.c_double(tol),
           C.c_int32(int(bool(demote_neighbours))), C.byref(handle), p(counts))
    try:
        nref_e = int(e2n.shape[0])
        cell_map, edge_map = np.empty(geo.ncell, np.int64), np.empty(nref_e, np.int64)
        edge_sigma, edge_inv = np.empty(nref_e, np.int8), np.empty(nedge_i, np.int64)
        sigma_int = np.empty(nedge_i, np.int8)
        reg_cell, reg_edge, reg_g1 = np.empty(ncell_i, np.uint8), np.empty(nedge_i, np.uint8), np.empty(ncell_i, np.uint8)
        comps = []
        for k, ncomp in enumerate((3, 1, 3)):
            nr, nnz = int(counts[2 * k]), int(counts[2 * k + 1])
            coef = np.empty((nnz, 3) if ncomp == 3 else (nnz,), np.float64)
            comps.append(Composite(np.empty(nr, np.int32), np.empty(nr + 1, np.int32), np.empty(nnz, np.int32), coef))
        arrs = [cell_map, edge_map, edge_sigma, edge_inv, sigma_int, reg_cell, reg_edge, reg_g1]
        for c in comps:
            arrs += [c.rows, c.ptr, c.col, c.coef]
        N.call("pyib_layout_fetch", handle, *[p(a) for a in arrs])
    finally:
        N.call("pyib_layout_free", handle)
    lay = StructuredLayout(levels, ncell_i, nedge_i, cell_map, edge_map, edge_sigma, edge_inv, sigma_int,
                           reg_cell, reg_edge, reg_g1, comps[0], comps[1], comps[2])
    lay.stats.update(regular_cells=int(counts[6]), regular_g1_cells=int(counts[7]), regular_edges=int(counts[8]))
    lay.stats.update(irregular_cells=int(counts[2]), irregular_g1_cells=int(counts[0]), irregular_edges=int(counts[4]),
                     g1_nnz=int(counts[1]), v_nnz=int(counts[3]), g3_nnz=int(counts[5]))
    return lay


def build_layout_numpy(mesh, geo, tables, margin=1, tol=1e-13, demote_neighbours=False):
    """NumPy statement of ``build_layout`` (the construction the C++ builder follows; the CPU
    tests compare the two array for array).  ``tol`` is the relative tolerance of the
    closed-form coefficient check (0: bitwise).  The default keeps rows whose
    geometry differs from the closed form only by the rounding of non-dyadic
    node coordinates (~1e-16 relative).  ``margin`` cells next to the box and the
    hole are never regular (the geometry check demotes the first layer anyway; one
    cell keeps every stencil read inside its level block).  ``demote_neighbours``
    also demotes the neighbours of geometrically irregular cells (not needed for
    exactness: the V stencil reads its neighbours' values, whichever pass wrote them)."""
    levels, ncell_i, nedge_i = _levels(mesh)
    N = len(levels)
    # ---- cells: reference layer order = C order of the live cells of each level
    cell_map = np.empty(geo.ncell, np.int64)
    start = 0
    for L, lv in enumerate(levels):
        live = np.flatnonzero(_cells_of_level(lv).ravel())
        n = int(mesh.layer_cell_counts[L])
        if live.shape[0] != n:
            raise ValueError("layer cell count does not match its box")
        cell_map[start:start + n] = lv.cell_off + live
        start += n
    # ---- edges: level from length, direction from the tangent, slot from the lower node
    nodes, E2N = mesh.nodes, mesh.edges_to_nodes
    te = geo.te
    d = np.argmax(np.abs(te), axis=1)
    length = np.abs(te[np.arange(te.shape[0]), d])
    h0 = levels[0].h
    lev = np.rint(np.log2(length / h0)).astype(np.int64)
    if lev.min() < 0 or lev.max() >= N or not np.allclose(length, h0 * 2.0 ** lev, rtol=1e-9):
        raise ValueError("edge length is not a level spacing")
    sigma = np.where(te[np.arange(te.shape[0]), d] > 0, 1, -1).astype(np.int8)
    lower = np.where((sigma > 0)[:, None], nodes[E2N[:, 0]], nodes[E2N[:, 1]])
    edge_map = np.empty(te.shape[0], np.int64)
    for L, lv in enumerate(levels):
        sel = lev == L
        ijk = np.rint((lower[sel] - np.asarray(lv.lo)) / lv.h).astype(np.int64)
        dd = d[sel]
        out = np.empty(int(sel.sum()), np.int64)
        for a in range(3):
            s = dd == a
            shp = _edge_shape(lv, a)
            q = ijk[s]
            if (q < 0).any() or (q >= np.asarray(shp)).any():
                raise ValueError("edge outside its level box")
            out[s] = lv.e_off[a] + (q[:, 0] * shp[1] + q[:, 1]) * shp[2] + q[:, 2]
        edge_map[sel] = out
    if np.unique(edge_map).shape[0] != edge_map.shape[0]:
        raise ValueError("two edges map to one slot")
    edge_inv = np.full(nedge_i, -1, np.int64)
    edge_inv[edge_map] = np.arange(edge_map.shape[0])
    sigma_int = np.ones(nedge_i, np.int8)
    sigma_int[edge_map] = sigma
    lay = StructuredLayout(levels, ncell_i, nedge_i, cell_map, edge_map, sigma, edge_inv, sigma_int,
                           np.zeros(ncell_i, np.uint8), np.zeros(nedge_i, np.uint8))
    _classify(lay, margin)
    _verify_regular(lay, mesh, geo, tables, tol, demote_neighbours, margin)
    _composites(lay, geo, tables)
    return lay


def _reg_cells_level(lv, m):
    r = np.zeros((lv.nx, lv.ny, lv.nz), bool)
    r[m:lv.nx - m, m:lv.ny - m, m:lv.nz - m] = True
    a, b = lv.hole_lo, lv.hole_hi
    if b[0] > a[0]:
        r[max(a[0] - m, 0):b[0] + m, max(a[1] - m, 0):b[1] + m, max(a[2] - m, 0):b[2] + m] = False
    return r


def _box_face_cells(lv, m):
    """Cells on the level box's boundary that lie at least m cells from the hole."""
    r = np.ones((lv.nx, lv.ny, lv.nz), bool)
    r[1:lv.nx - 1, 1:lv.ny - 1, 1:lv.nz - 1] = False
    a, b = lv.hole_lo, lv.hole_hi
    if b[0] > a[0]:
        r[max(a[0] - m, 0):b[0] + m, max(a[1] - m, 0):b[1] + m, max(a[2] - m, 0):b[2] + m] = False
    return r


def _reg_edges_level(lv, rc, d):
    """An edge is regular when its four adjacent cells are regular."""
    shp = _edge_shape(lv, d)
    out = np.zeros(shp, bool)
    o = [a for a in range(3) if a != d]
    # cells adjacent to edge (i,j,k) along d: offsets -1/0 on the two other axes
    pad = np.zeros(tuple(s + 2 for s in rc.shape), bool)
    pad[1:-1, 1:-1, 1:-1] = rc
    acc = np.ones(shp, bool)
    for s0 in (-1, 0):
        for s1 in (-1, 0):
            sl = [slice(1, 1 + shp[0]), slice(1, 1 + shp[1]), slice(1, 1 + shp[2])]
            for ax, sft in ((o[0], s0), (o[1], s1)):
                sl[ax] = slice(1 + sft, 1 + sft + shp[ax])
            acc &= pad[tuple(sl)]
    out[:] = acc
    return out


def _classify(lay, m):
    for lv in lay.levels:
        rc = _reg_cells_level(lv, m)
        lay.reg_cell[lv.cell_off:lv.cell_off + lv.ncell] = rc.ravel()
        for d in range(3):
            re = _reg_edges_level(lv, rc, d)
            n = re.size
            lay.reg_edge[lv.e_off[d]:lv.e_off[d] + n] = re.ravel()


def _close(a, b, tol, scale):
    """|a - b| <= tol*scale (bitwise equality when tol == 0)."""
    return np.abs(a - b) <= tol * scale if tol > 0 else a == b


def _verify_regular(lay, mesh, geo, tables, tol, demote_neighbours=True, margin=1):
    """Regularity per pass, from the mesh geometry (each pass only needs its own data):

    * G1 (cell <- edges): closed-form vol_ci, exactly six faces, and on each face this
      cell's rf (rf1 as first owner, rf2 as second) = +-h/2 e_d with four links;
    * V (cell <- cells): closed-form vol_ci and six internal same-level faces with AoL = h,
      at least ``margin`` cells from the box and the hole;
    * G3 (edge <- cells): an edge off the box boundary whose four adjacent cells are real
      and whose four faces (exactly four rot links) have closed-form rf1 and rf2.
    """
    ncell = geo.ncell
    lev_of_cell = np.empty(ncell, np.int64)
    for L, lv in enumerate(lay.levels):
        sel = (lay.cell_map >= lv.cell_off) & (lay.cell_map < lv.cell_off + lv.ncell)
        lev_of_cell[sel] = L
    h = np.array([lv.h for lv in lay.levels])[lev_of_cell]
    ok_vol = _close(geo.vol_ci, 1.0 / h ** 3, tol, 1.0 / h ** 3)
    F2C = geo.F2C
    hf = h[F2C[:, 0]]
    dn = np.argmax(np.abs(geo.nf), axis=1)
    e_d = np.zeros((geo.nface, 3))
    e_d[np.arange(geo.nface), dn] = 1.0
    half = 0.5 * hf[:, None] * e_d
    sgn = np.sign(geo.rf1[np.arange(geo.nface), dn])
    links4 = np.diff(tables.curl_ptr) == 4
    same_h = geo.m1 & (h[np.maximum(F2C[:, 1], 0)] == hf)
    rf1_ok = np.all(_close(np.abs(geo.rf1), half, tol, hf[:, None]), axis=1) & (sgn != 0)
    rf2_ok = same_h & np.all(_close(np.abs(geo.rf2), half, tol, hf[:, None]), axis=1)
    ok_face = rf1_ok & rf2_ok & links4 & _close(np.abs(geo.AoL), hf, tol, hf) & _close(geo.Af, hf * hf, tol, hf * hf)
    # per (cell, face) entries of the cell-face table
    cnt = np.diff(tables.cf_ptr)
    own = np.repeat(np.arange(ncell), cnt)
    pos = np.arange(tables.cf_face.shape[0]) - np.repeat(tables.cf_ptr[:-1], cnt)
    first = pos < np.repeat(tables.cf_mid - tables.cf_ptr[:-1], cnt)
    f = tables.cf_face
    g1_ent = np.where(first, rf1_ok[f], rf2_ok[f]) & links4[f]
    g1_ok = ok_vol & (cnt == 6) & (np.bincount(own, weights=(~g1_ent).astype(np.float64), minlength=ncell) == 0)
    # V reads only vol_ci, AoL and the six same-level neighbours
    v_face = same_h & _close(np.abs(geo.AoL), hf, tol, hf)
    v_ok = ok_vol & (cnt == 6) & (np.bincount(own, weights=(~v_face[f]).astype(np.float64), minlength=ncell) == 0)
    real = np.zeros(lay.ncell, bool)
    real[lay.cell_map] = True
    reg_g1 = np.zeros(lay.ncell, bool)
    reg_g1[lay.cell_map[g1_ok]] = True
    reg_v = lay.reg_cell.astype(bool)           # structural (margin) part from _classify
    v_geo = np.zeros(lay.ncell, bool)
    v_geo[lay.cell_map[v_ok]] = True
    reg_v &= v_geo & real
    # box-face V rows: cells of the coarsest level on the domain boundary (at least margin cells from
    # the hole) whose faces inside the box are same-level faces with AoL = h and whose other faces are
    # boundary faces (zero ghost) with AoL = 2h; the stencil gives those faces 2 kA (reg_cell bit 1)
    lc = lay.levels[-1]
    bface = ~np.asarray(geo.m1) & _close(geo.AoL, 2.0 * hf, tol, hf)
    vb_ent = np.where(tables.cf_other < ncell, v_face[f], bface[f])
    vb_ok = ok_vol & (cnt == 6) & (np.bincount(own, weights=(~vb_ent).astype(np.float64), minlength=ncell) == 0)
    reg_vb = np.zeros(lay.ncell, bool)
    reg_vb[lay.cell_map[vb_ok]] = True
    box = np.zeros(lay.ncell, bool)
    box[lc.cell_off:lc.cell_off + lc.ncell] = _box_face_cells(lc, margin).ravel()
    reg_vb &= box & real & ~reg_v
    reg_v |= reg_vb
    if demote_neighbours:
        bad = real & ~reg_v
        nb = np.zeros(ncell, bool)
        inner = tables.cf_other < ncell
        cinv = np.full(lay.ncell, -1, np.int64)
        cinv[lay.cell_map] = np.arange(ncell)
        badr = bad[lay.cell_map]
        nb[own[inner][badr[tables.cf_other[inner]]]] = True
        reg_v[lay.cell_map[nb]] = False
        reg_vb &= reg_v
    lay.reg_g1 = reg_g1.astype(np.uint8)
    lay.reg_cell = reg_v.astype(np.uint8) | (reg_vb.astype(np.uint8) << 1)
    # G3: structural (off the box boundary, four real adjacent cells) and geometric
    ecnt = np.diff(tables.rot_ptr)
    eown = np.repeat(np.arange(geo.nedge), ecnt)
    fidx = tables.rot_idx.astype(np.int64)
    fidx = np.where(fidx >= 0, fidx, -fidx - 1)
    f_g3 = rf1_ok & rf2_ok & links4
    g3_ok = (ecnt == 4) & (np.bincount(eown, weights=(~f_g3[fidx]).astype(np.float64), minlength=geo.nedge) == 0)
    reg_e = np.zeros(lay.nedge, bool)
    reg_e[lay.edge_map[g3_ok]] = True
    for lv in lay.levels:
        rc = real[lv.cell_off:lv.cell_off + lv.ncell].reshape(lv.nx, lv.ny, lv.nz)
        for d in range(3):
            st = _reg_edges_level(lv, rc, d)        # the four adjacent cells exist (real, in the box)
            seg = slice(lv.e_off[d], lv.e_off[d] + st.size)
            reg_e[seg] &= st.ravel()
    reg_e[lay.edge_inv < 0] = False
    lay.reg_edge = reg_e.astype(np.uint8)
    lay.stats.update(regular_cells=int(reg_v.sum()), regular_g1_cells=int(reg_g1.sum()),
                     regular_edges=int(reg_e.sum()))


def expand_ranges(starts, counts):
    """Concatenation of arange(starts[i], starts[i]+counts[i]) for all i (vectorised)."""
    starts = np.asarray(starts, np.int64)
    counts = np.asarray(counts, np.int64)
    total = int(counts.sum())
    if total == 0:
        return np.zeros(0, np.int64)
    first = np.cumsum(counts) - counts
    return np.arange(total, dtype=np.int64) - np.repeat(first, counts) + np.repeat(starts, counts)


def _csr_rows(rows_of_entry, n_rows_total, cols, coefs, row_ids):
    """Sum duplicate (row, col) entries and pack the rows listed in row_ids (internal ids)."""
    key = rows_of_entry.astype(np.int64) * (np.int64(1) << 32) + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key, coefs = key[order], coefs[order]
    uniq, start = np.unique(key, return_index=True)
    summed = np.add.reduceat(coefs, start, axis=0) if coefs.shape[0] else coefs
    r = (uniq >> 32).astype(np.int64)
    c = (uniq & 0xFFFFFFFF).astype(np.int64)
    counts = np.bincount(r, minlength=n_rows_total)
    full_ptr = np.zeros(n_rows_total + 1, np.int64)
    np.cumsum(counts, out=full_ptr[1:])
    sel_counts = counts[row_ids]
    ptr = np.zeros(row_ids.shape[0] + 1, np.int64)
    np.cumsum(sel_counts, out=ptr[1:])
    take = expand_ranges(full_ptr[row_ids], sel_counts)
    return ptr.astype(np.int32), c[take].astype(np.int32), summed[take]


def _composites(lay, geo, t):
    ncell = geo.ncell
    cmap, emap, sig = lay.cell_map, lay.edge_map, lay.edge_sigma.astype(np.float64)
    real_cell = np.zeros(lay.ncell, bool)
    real_cell[cmap] = True
    irr_g1_int = np.flatnonzero(real_cell & (lay.reg_g1 == 0))
    irr_v_int = np.flatnonzero(real_cell & (lay.reg_cell == 0))
    irr_cells_int = np.union1d(irr_g1_int, irr_v_int)
    cinv = np.full(lay.ncell, -1, np.int64)
    cinv[cmap] = np.arange(ncell)
    irr_ref = cinv[irr_cells_int]
    # cell-face entries of the irregular cells
    cnt = np.diff(t.cf_ptr)
    sel = np.zeros(ncell, bool)
    sel[irr_ref] = True
    own = np.repeat(np.arange(ncell), cnt)
    first = np.zeros(t.cf_face.shape[0], bool)
    pos = np.arange(t.cf_face.shape[0]) - np.repeat(t.cf_ptr[:-1], cnt)
    first[pos < np.repeat(t.cf_mid - t.cf_ptr[:-1], cnt)] = True
    ent = np.flatnonzero(sel[own])
    f_e, own_e, first_e, oth_e = t.cf_face[ent], own[ent], first[ent], t.cf_other[ent]
    # --- G1: cell <- edge
    rfc = np.where(first_e[:, None], geo.rf1[f_e], -geo.rf2[f_e]) * geo.vol_ci[own_e][:, None]
    lcnt = np.diff(t.curl_ptr)[f_e]
    rep = np.repeat(np.arange(f_e.shape[0]), lcnt)
    link = expand_ranges(t.curl_ptr[f_e], lcnt)
    sv = t.curl_idx[link].astype(np.int64)
    e_ref = np.where(sv >= 0, sv, -sv - 1)
    s_link = np.where(sv >= 0, 1.0, -1.0)
    coef = rfc[rep] * (s_link * sig[e_ref])[:, None]
    ptr, col, cf = _csr_rows(cmap[own_e][rep], lay.ncell, emap[e_ref], coef, irr_g1_int)
    lay.g1 = Composite(irr_g1_int.astype(np.int32), ptr, col, cf)
    # --- V (cell <- cell): off-diagonal vol_ci*AoL, diagonal -vol_ci*sum AoL
    aol = geo.AoL[f_e] * geo.vol_ci[own_e]
    inner = oth_e < ncell
    rows = np.concatenate([cmap[own_e][inner], cmap[own_e]])
    cols = np.concatenate([cmap[oth_e[inner]], cmap[own_e]])
    vals = np.concatenate([aol[inner], -aol])
    ptr, col, cf = _csr_rows(rows, lay.ncell, cols, vals, irr_v_int)
    lay.vis = Composite(irr_v_int.astype(np.int32), ptr, col, cf)
    # --- G3: edge <- cell, for real irregular edges
    real_edge = lay.edge_inv >= 0
    irr_edges_int = np.flatnonzero(real_edge & (lay.reg_edge == 0))
    er = lay.edge_inv[irr_edges_int]
    rcnt = np.diff(t.rot_ptr)[er]
    rep = np.repeat(np.arange(er.shape[0]), rcnt)
    link = expand_ranges(t.rot_ptr[er], rcnt)
    sv = t.rot_idx[link].astype(np.int64)
    f_ref = np.where(sv >= 0, sv, -sv - 1)
    s_link = np.where(sv >= 0, 1.0, -1.0) * sig[er[rep]]
    c1, c2 = geo.F2C[f_ref, 0], geo.F2C[f_ref, 1]
    rows = np.concatenate([irr_edges_int[rep], irr_edges_int[rep][c2 >= 0]])
    cols = np.concatenate([cmap[c1], cmap[c2[c2 >= 0]]])
    vals = np.concatenate([geo.rf1[f_ref] * s_link[:, None], -geo.rf2[f_ref][c2 >= 0] * s_link[c2 >= 0][:, None]])
    ptr, col, cf = _csr_rows(rows, lay.nedge, cols, vals, irr_edges_int)
    lay.g3 = Composite(irr_edges_int.astype(np.int32), ptr, col, cf)
    lay.stats.update(irregular_cells=int(irr_v_int.shape[0]), irregular_g1_cells=int(irr_g1_int.shape[0]),
                     irregular_edges=int(irr_edges_int.shape[0]),
                     g1_nnz=int(lay.g1.col.shape[0]), v_nnz=int(lay.vis.col.shape[0]), g3_nnz=int(lay.g3.col.shape[0]))


SLICE = 64   # rows per slice of the sliced-ELL composites: one wavefront


@dataclass
class SellComposite:
    """Composite rows in sliced ELL (slices of SLICE rows, column-major inside a slice)."""
    rows: np.ndarray            # (nslice*SLICE,) int32 internal row id, -1 on padding lanes
    width: np.ndarray           # (nslice,) int32 entries per lane
    sptr: np.ndarray            # (nslice,) int64 first entry of the slice
    col: np.ndarray             # int32, 0 on padding
    coef: np.ndarray            # float64 (entries,) or (entries, 3), 0 on padding


def to_sell(comp, C=SLICE, sigma=8 * SLICE):
    """CSR composite -> sliced ELL.  Inside windows of ``sigma`` consecutive rows (rows come
    in internal, i.e. spatial, order) the rows are ordered by decreasing length, which keeps
    the padding small while neighbouring lanes still gather from neighbouring cells; each
    lane keeps its row's entries in CSR order, so the per-row sums are unchanged."""
    n = int(comp.rows.shape[0])
    width = np.diff(comp.ptr).astype(np.int64)
    window = np.arange(n, dtype=np.int64) // max(int(sigma), 1)
    order = np.lexsort((-width, window))
    nslice = (n + C - 1) // C
    w_sorted = np.zeros(nslice * C, np.int64)
    w_sorted[:n] = width[order]
    sw = w_sorted.reshape(nslice, C).max(axis=1) if nslice else np.zeros(0, np.int64)
    sptr = np.zeros(nslice, np.int64)
    if nslice:
        np.cumsum(sw[:-1] * C, out=sptr[1:])
    total = int((sw * C).sum())
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    ent_row = np.repeat(np.arange(n), width)
    ent_j = np.arange(int(width.sum())) - np.repeat(comp.ptr[:-1].astype(np.int64), width)
    p = pos[ent_row]
    dst = sptr[p // C] + ent_j * C + p % C
    col = np.zeros(total, np.int32)
    col[dst] = comp.col
    coef = np.zeros((total,) + comp.coef.shape[1:], comp.coef.dtype)
    coef[dst] = comp.coef
    rows = np.full(nslice * C, -1, np.int32)
    rows[:n] = comp.rows[order]
    return SellComposite(rows, sw.astype(np.int32), sptr, col, coef)


def composite_rows(comp, sel):
    """The composite made of rows ``sel`` (indices into comp.rows) of ``comp``, each row's
    entries in their original order."""
    sel = np.asarray(sel, dtype=np.int64)
    ptr = np.asarray(comp.ptr, dtype=np.int64)
    width = ptr[sel + 1] - ptr[sel]
    new_ptr = np.zeros(sel.shape[0] + 1, np.int64)
    np.cumsum(width, out=new_ptr[1:])
    ent = np.repeat(ptr[sel], width) + (np.arange(int(new_ptr[-1])) - np.repeat(new_ptr[:-1], width))
    return Composite(np.asarray(comp.rows)[sel], new_ptr, np.asarray(comp.col)[ent], np.asarray(comp.coef)[ent])


def g1v_flags(lay):
    """Per internal cell, the flags of the fused G1+V pass: bit 0 = regular G1 row (x =
    interp curl p formed by the stencil), bit 1 = that x must also be stored, because a
    composite V row reads it (as its own row or a column), bit 2 = regular V row, bit 3 = x
    comes from memory (a real cell whose G1 row is composite: the composite G1 pass wrote it),
    bit 4 = a box-face V row (its faces that leave the box take twice the stencil coefficient).
    The pass forms x in LDS, stores only the bit-1 cells, and takes x = 0 on the cells with
    neither bit 0 nor bit 3 (the hole, whose x no stencil row reads)."""
    keep = np.zeros(lay.ncell, bool)
    keep[np.asarray(lay.vis.rows, dtype=np.int64)] = True
    keep[np.asarray(lay.vis.col, dtype=np.int64)] = True
    reg = lay.reg_g1.astype(bool)
    regv = (lay.reg_cell & 1).astype(bool)
    boxv = (lay.reg_cell & 2).astype(bool)
    real = np.zeros(lay.ncell, bool)
    real[np.asarray(lay.cell_map, dtype=np.int64)] = True
    return (reg.astype(np.uint8) | ((reg & keep).astype(np.uint8) << 1)
            | (regv.astype(np.uint8) << 2) | ((real & ~reg).astype(np.uint8) << 3)
            | (boxv.astype(np.uint8) << 4)).astype(np.uint8)


def row_volumes(lay, geo, rows):
    """Cell volume 1/vol_ci of each internal cell row (0 for -1 padding and for cells with
    vol_ci <= 0).  The composite V kernel weighs each row's x.(V x) by it to form its share
    of p.A p: A = rot integr V interp curl with integr = interp^T diag(1/vol_ci)
    (ib_sparse.py:29-50), so p.A p = sum_cells vol x.(V x) with x = interp curl p."""
    vol = np.zeros(lay.ncell)
    vci = np.asarray(geo.vol_ci, dtype=np.float64)
    pos = vci > 0
    vol[lay.cell_map[pos]] = 1.0 / vci[pos]
    rows = np.asarray(rows)
    return np.where(rows >= 0, vol[np.maximum(rows, 0)], 0.0)


def sell_apply(sc, x_gather, C=SLICE):
    """Host emulation of a sliced-ELL row sum: out[row] = sum_j coef * x_gather(col), summed
    sequentially in entry order as the kernels do (padding entries add +0)."""
    out = {}
    for s in range(sc.width.shape[0]):
        for l in range(C):
            r = int(sc.rows[s * C + l])
            if r < 0:
                continue
            idx = sc.sptr[s] + np.arange(int(sc.width[s])) * C + l
            terms = sc.coef[idx] * x_gather(sc.col[idx])
            out[r] = np.cumsum(terms, axis=0)[-1] if terms.shape[0] else np.zeros(terms.shape[1:])
    return out


def dir_mask_internal(lay, dir_mask_ref):
    """Dirichlet mask in internal numbering; empty slots count as masked."""
    m = np.ones(lay.nedge, np.uint8)
    m[lay.edge_map] = np.asarray(dir_mask_ref, bool).astype(np.uint8)
    return m


def edge_classes(lay, dir_int, prec_int):
    """Classes of the Jacobi preconditioner over the internal edge slots.

    ``prec_int`` is the inverse diagonal of ``diag_preconditioner`` (ib_driver_np.py:78-88) in
    internal numbering (0 on Dirichlet rows and empty slots).  On the edges of a level whose four
    faces are whole h x h squares it is the same number (the sum of 1/h over the four faces,
    inverted), so the PCG iteration can take it from a per-level table.  The table value of a
    level is its most frequent value over the live slots, and an edge is given bit 2 only where
    its own value equals it bitwise, so the table never changes a result.

    Returns (cls uint8 (nedge,), prec_level (MAX_LEVELS,), prec_level_g3 (MAX_LEVELS,)):
    cls bit 0 = regular G3 row, bit 1 = live (neither empty nor Dirichlet), bit 2 = prec equals
    prec_level of its level, bits 4-6 = level; prec_level_g3[l] = prec_level[l] when every
    regular G3 row of level l has bit 2, else 0."""
    prec_int = np.asarray(prec_int, np.float64)
    dir_int = np.asarray(dir_int)
    if prec_int.shape != (lay.nedge,) or dir_int.shape != (lay.nedge,):
        raise ValueError("edge_classes needs the internal Dirichlet mask and preconditioner")
    reg = lay.reg_edge.astype(bool)
    live = dir_int == 0
    cls = reg.astype(np.uint8) | (live.astype(np.uint8) << 1)
    plev = np.zeros(MAX_LEVELS)
    pg3 = np.zeros(MAX_LEVELS)
    bounds = [lv.e_off[0] for lv in lay.levels] + [lay.nedge]
    for L in range(len(lay.levels)):
        a, b = bounds[L], bounds[L + 1]
        cls[a:b] |= np.uint8(L << 4)
        seg, lv_live = prec_int[a:b], live[a:b]
        if not lv_live.any():
            continue
        vals, counts = np.unique(seg[lv_live], return_counts=True)
        c = float(vals[np.argmax(counts)])
        closed = lv_live & (seg == c)
        cls[a:b] |= closed.astype(np.uint8) << 2
        plev[L] = c
        r = reg[a:b]
        if c != 0.0 and bool(np.all(closed[r])):
            pg3[L] = c
    return cls, plev, pg3


def to_internal_edges(lay, x_ref):
    out = np.zeros(lay.nedge)
    out[lay.edge_map] = lay.edge_sigma * np.asarray(x_ref)
    return out


def to_reference_edges(lay, x_int):
    return lay.edge_sigma * np.asarray(x_int)[lay.edge_map]
