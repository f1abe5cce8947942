"""Nested 2:1 Cartesian hex mesh in PyIB's global numbering (host setup).

Builds the same arrays as the reference ``make_nested_mesh``
(``src/pyib/make_euler_mesh.py:330-463``): node ids assigned outermost layer
first, cells layer by layer in C order, faces and edges numbered by the
lexicographic order of their sorted node ids, then the hanging-node
finalisation of ``make_euler_mesh.py:787-950`` (fine interface faces take the
coarse cell as second owner, coarse interface faces are removed, each coarse
interface edge is collapsed onto its two fine halves with the second half
recorded in ``link_fe_insert``).  Every step here is vectorised (packed
integer keys + sort/searchsorted instead of Python dictionaries), so the
40M-edge meshes of the 3-level D32/D40 oscillating-sphere cases build in
seconds to minutes.  The ordering is bit-identical to the reference, which the
sha256 layout fingerprint of the reference D32/D40 meshes pins.

Only refinement ratio 2 is supported: the geometry correction of the solver
(``ib_core_np.py:123-147``) is written for ratio 2.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

CORNERS = np.array([(1, 0, 1), (1, 0, 0), (0, 0, 1), (0, 0, 0),
                    (1, 1, 1), (1, 1, 0), (0, 1, 1), (0, 1, 0)], dtype=np.int64)
HEX_FACES = np.array([[3, 1, 0, 2], [7, 5, 4, 6], [3, 1, 5, 7],
                      [2, 0, 4, 6], [3, 7, 6, 2], [1, 5, 4, 0]], dtype=np.int64)
FACE_EDGES = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], dtype=np.int64)
DIAMETER = 0.5


@dataclass(frozen=True)
class BoxSpec:
    """Axis-aligned box (``make_euler_mesh.py:134-161``)."""
    center: tuple
    size: tuple

    @classmethod
    def from_bounds(cls, xl, xh, yl, yh, zl, zh):
        if not (xl < xh and yl < yh and zl < zh):
            raise ValueError("from_bounds requires lo < hi in each axis")
        return cls(((xl + xh) / 2.0, (yl + yh) / 2.0, (zl + zh) / 2.0), (xh - xl, yh - yl, zh - zl))

    @property
    def min(self):
        return np.array(self.center, dtype=np.float64) - np.array(self.size, dtype=np.float64) / 2.0

    @property
    def max(self):
        return np.array(self.center, dtype=np.float64) + np.array(self.size, dtype=np.float64) / 2.0


@dataclass(frozen=True)
class CombinedMesh:
    """Same fields (and fingerprint order) as ``make_euler_mesh.py:164-195``."""
    nodes: np.ndarray
    cells: np.ndarray
    n_layers: int
    layer_cell_counts: tuple
    layer_node_counts: tuple
    refinement_ratio: int
    boundary_faces: np.ndarray
    boundary_fam: np.ndarray
    pos_ref: np.ndarray
    dx_uniform: np.ndarray
    index_cell: np.ndarray
    faces_to_nodes: np.ndarray
    faces_to_cells: np.ndarray
    edges_to_nodes: np.ndarray
    faces_to_edges: np.ndarray
    bc_node: np.ndarray
    bc_edge: np.ndarray
    bc_face: np.ndarray
    bc_cell: np.ndarray
    interface_pairs: np.ndarray
    link_fe_insert: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))
    # layer geometry kept for the kernels' structured view (not part of the fingerprint)
    layer_spacing: tuple = ()
    layer_shape: tuple = ()
    layer_min: tuple = ()
    layer_hole_lo: tuple = ()


def _layers(h, boxes):
    """Spacing, cell counts and hole offsets per layer (``make_euler_mesh.py:234-290``)."""
    h = np.atleast_1d(np.asarray(h, dtype=np.float64))
    if h.size == 1:
        h = np.array([float(h.flat[0])] * 3, dtype=np.float64)
    dx = [h * (2 ** k) for k in range(len(boxes))]
    shape = []
    for k, b in enumerate(boxes):
        size = np.asarray(b.size, dtype=np.float64)
        n = np.round(size / dx[k]).astype(np.int64)
        if not np.allclose(size, n * dx[k], atol=1e-8):
            raise ValueError(f"layer {k} box size is not a multiple of dx={tuple(dx[k])}")
        shape.append(tuple(int(v) for v in n))
    hole = [np.zeros(3, np.int64)]
    for k in range(1, len(boxes)):
        inner, outer = boxes[k - 1], boxes[k]
        if (inner.min <= outer.min + 1e-9).any() or (inner.max >= outer.max - 1e-9).any():
            raise ValueError(f"layer {k-1} box must lie strictly inside layer {k} box")
        off = inner.min - outer.min
        oi = np.round(off / dx[k]).astype(np.int64)
        if not np.allclose(off, oi * dx[k], atol=1e-8):
            raise ValueError(f"layer {k-1} corners are not aligned with layer {k}'s grid")
        if any(shape[k - 1][a] % 2 for a in range(3)):
            raise ValueError(f"layer {k-1} cell counts must be even for 2:1 nesting")
        hole.append(oi)
    return dx, shape, hole


def _pack2(a, b):
    return (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)


def _face_keys(quads):
    s = np.sort(quads, axis=1)
    return _pack2(s[:, 0], s[:, 1]), _pack2(s[:, 2], s[:, 3])


def _lookup2(hi_sorted, lo_sorted, hi, lo):
    """Row index of (hi, lo) in the lexicographically sorted (hi_sorted, lo_sorted); -1 if absent."""
    left = np.searchsorted(hi_sorted, hi, side="left")
    right = np.searchsorted(hi_sorted, hi, side="right")
    out = np.full(hi.shape[0], -1, np.int64)
    # faces sharing the first two node ids are rare (<= 4); scan the short ranges
    width = right - left
    for j in range(int(width.max()) if width.size else 0):
        idx = left + j
        ok = (j < width) & (out < 0)
        idx_c = np.minimum(idx, hi_sorted.shape[0] - 1)
        hit = ok & (lo_sorted[idx_c] == lo)
        out[hit] = idx[hit]
    return out


def make_nested_mesh(inner_spacing, boxes, refinement_ratio=2, *, verbose=False) -> CombinedMesh:
    if int(refinement_ratio) != 2:
        raise ValueError("only refinement_ratio=2 is supported")
    dx, shape, hole = _layers(inner_spacing, boxes)
    N = len(boxes)
    nshape = [(s[0] + 1, s[1] + 1, s[2] + 1) for s in shape]
    lo = [b.min for b in boxes]

    def hole_mask_nodes(k):
        m = np.zeros(nshape[k], bool)
        if k > 0:
            hn = np.asarray(shape[k - 1]) // 2
            a = hole[k]
            m[a[0] + 1:a[0] + hn[0], a[1] + 1:a[1] + hn[1], a[2] + 1:a[2] + hn[2]] = True
        return m

    ids = [None] * N
    nnode = [0] * N
    nxt = 0
    for k in range(N - 1, -1, -1):
        ids[k] = np.full(nshape[k], -1, np.int64)
        if k < N - 1:
            o = hole[k + 1]
            ids[k][::2, ::2, ::2] = ids[k + 1][o[0]:o[0] + nshape[k][0] // 2 + 1,
                                               o[1]:o[1] + nshape[k][1] // 2 + 1,
                                               o[2]:o[2] + nshape[k][2] // 2 + 1]
        need = (ids[k] < 0) & ~hole_mask_nodes(k)
        n = int(need.sum())
        ids[k][need] = np.arange(nxt, nxt + n)
        nnode[k] = n
        nxt += n
    nodes = np.zeros((nxt, 3), np.float64)
    for k in range(N):
        a, b, c = np.where(ids[k] >= 0)
        nodes[ids[k][a, b, c]] = np.column_stack([lo[k][0] + a * dx[k][0], lo[k][1] + b * dx[k][1],
                                                  lo[k][2] + c * dx[k][2]])
    per_layer = []
    for k in range(N):
        mask = np.ones(shape[k], bool)
        if k > 0:
            hn = np.asarray(shape[k - 1]) // 2
            a = hole[k]
            mask[a[0]:a[0] + hn[0], a[1]:a[1] + hn[1], a[2]:a[2] + hn[2]] = False
        g = ids[k]
        nx, ny, nz = shape[k]
        corners = np.stack([g[p:p + nx, q:q + ny, t:t + nz] for p, q, t in CORNERS], axis=-1)
        per_layer.append(corners[mask & (corners >= 0).all(axis=-1)])
        del corners
    layer_cells = [int(c.shape[0]) for c in per_layer]
    cells = np.concatenate(per_layer, axis=0)
    del per_layer
    ncell = cells.shape[0]

    # outer boundary quads and families (make_euler_mesh.py:481-531)
    g = ids[N - 1]
    X, Y, Z = shape[N - 1]
    q = lambda a, b, c, d: np.stack([a, b, c, d], axis=-1).reshape(-1, 4)
    fams = [
        q(g[0, 0:Y, 1:Z + 1], g[0, 1:Y + 1, 1:Z + 1], g[0, 1:Y + 1, 0:Z], g[0, 0:Y, 0:Z]),
        q(g[X, 1:Y + 1, 1:Z + 1], g[X, 0:Y, 1:Z + 1], g[X, 0:Y, 0:Z], g[X, 1:Y + 1, 0:Z]),
        q(g[0:X, 0:Y, Z], g[1:X + 1, 0:Y, Z], g[1:X + 1, 1:Y + 1, Z], g[0:X, 1:Y + 1, Z]),
        q(g[0:X, 0:Y, 0], g[0:X, 1:Y + 1, 0], g[1:X + 1, 1:Y + 1, 0], g[1:X + 1, 0:Y, 0]),
        q(g[0:X, Y, 1:Z + 1], g[1:X + 1, Y, 1:Z + 1], g[1:X + 1, Y, 0:Z], g[0:X, Y, 0:Z]),
        q(g[1:X + 1, 0, 1:Z + 1], g[0:X, 0, 1:Z + 1], g[0:X, 0, 0:Z], g[1:X + 1, 0, 0:Z]),
    ]
    bfaces = np.concatenate(fams, axis=0).astype(np.int64, copy=False)
    bfam = np.concatenate([np.full(f.shape[0], t, np.int64) for t, f in enumerate(fams, 1)])

    # faces: lexicographic order of sorted node ids; stable => c1 is the smaller cell id
    allf = cells[:, HEX_FACES].reshape(-1, 4)
    khi, klo = _face_keys(allf)
    order = np.lexsort((klo, khi))
    del klo
    shi = khi[order]
    del khi
    slo = _face_keys(allf[order])[1]
    new = np.empty(order.shape[0], bool)
    new[0] = True
    new[1:] = (shi[1:] != shi[:-1]) | (slo[1:] != slo[:-1])
    start = np.flatnonzero(new)
    nface = start.shape[0]
    cnt = np.diff(np.append(start, order.shape[0]))
    if (cnt > 2).any():
        raise ValueError("non-manifold: a face is adjacent to >2 cells")
    f2n = allf[order[start]]
    owner = order // 6
    f2c = np.full((nface, 2), -1, np.int64)
    f2c[:, 0] = owner[start]
    two = cnt == 2
    f2c[two, 1] = owner[start[two] + 1]
    uhi, ulo = shi[start], slo[start]
    del shi, slo, order, owner, allf, new
    bhi, blo = _face_keys(bfaces)
    fid = _lookup2(uhi, ulo, bhi, blo)
    if (fid < 0).any() or (cnt[fid] != 1).any():
        raise ValueError("boundary quad does not match a boundary face")
    bc_face = np.zeros(nface, np.int64)
    bc_face[fid] = bfam

    # edges: lexicographic order of (min node, max node)
    fe = np.sort(f2n[:, FACE_EDGES].reshape(-1, 2), axis=1)
    ekey = _pack2(fe[:, 0], fe[:, 1])
    del fe
    ukeys, inv = np.unique(ekey, return_inverse=True)
    del ekey
    e2n = np.column_stack([(ukeys >> np.uint64(32)).astype(np.int64),
                           (ukeys & np.uint64(0xFFFFFFFF)).astype(np.int64)])
    f2e = inv.reshape(nface, 4).astype(np.int64)
    del inv
    bc_node = np.zeros(nxt, np.int64)
    bc_edge = np.zeros(e2n.shape[0], np.int64)
    bm = bc_face != 0
    np.maximum.at(bc_node, f2n[bm].ravel(), np.repeat(bc_face[bm], 4))
    np.maximum.at(bc_edge, f2e[bm].ravel(), np.repeat(bc_face[bm], 4))
    bc_cell = np.zeros(ncell, np.int64)

    # interface pairs (make_euler_mesh.py:606-709), same emission order
    pairs = []
    sides = ((0, False, 4, 5), (0, True, 5, 4), (1, False, 0, 1),
             (1, True, 1, 0), (2, False, 2, 3), (2, True, 3, 2))
    for k in range(N - 1):
        a_lo = hole[k + 1]
        nk = np.asarray(shape[k], np.int64)
        a_hi = a_lo + nk // 2
        alive_c = np.ones(shape[k + 1], bool)
        alive_c[a_lo[0]:a_hi[0], a_lo[1]:a_hi[1], a_lo[2]:a_hi[2]] = False
        cid_c = np.full(shape[k + 1], -1, np.int64)
        off = sum(layer_cells[:k + 1])
        cid_c[alive_c] = np.arange(off, off + int(alive_c.sum()))
        alive_f = np.ones(shape[k], bool)
        if k > 0:
            h0 = hole[k]
            hn = np.asarray(shape[k - 1]) // 2
            alive_f[h0[0]:h0[0] + hn[0], h0[1]:h0[1] + hn[1], h0[2]:h0[2] + hn[2]] = False
        cid_f = np.full(shape[k], -1, np.int64)
        off = sum(layer_cells[:k])
        cid_f[alive_f] = np.arange(off, off + int(alive_f.sum()))
        for axis, is_max, fine_lf, coarse_lf in sides:
            o0, o1 = [a for a in (0, 1, 2) if a != axis]
            c0, c1 = np.meshgrid(np.arange(a_lo[o0], a_hi[o0]), np.arange(a_lo[o1], a_hi[o1]), indexing="ij")
            c0, c1 = c0.ravel(), c1.ravel()
            ci = [None] * 3
            ci[axis] = np.full(c0.shape, a_hi[axis] if is_max else a_lo[axis] - 1)
            ci[o0], ci[o1] = c0, c1
            coarse = cid_c[ci[0], ci[1], ci[2]]
            if (coarse < 0).any():
                raise ValueError("interface coarse cell missing")
            chi, clo = _face_keys(cells[coarse][:, HEX_FACES[coarse_lf]])
            cols = [_lookup2(uhi, ulo, chi, clo)]
            for d0 in range(2):
                for d1 in range(2):
                    fi = [None] * 3
                    fi[axis] = np.full(c0.shape, nk[axis] - 1 if is_max else 0)
                    fi[o0] = (c0 - a_lo[o0]) * 2 + d0
                    fi[o1] = (c1 - a_lo[o1]) * 2 + d1
                    fc = cid_f[fi[0], fi[1], fi[2]]
                    fhi, flo = _face_keys(cells[fc][:, HEX_FACES[fine_lf]])
                    cols.append(_lookup2(uhi, ulo, fhi, flo))
            blk = np.column_stack(cols)
            if (blk < 0).any():
                raise ValueError("interface face lookup failed")
            pairs.append(blk)
    pairs = np.concatenate(pairs, axis=0) if pairs else np.zeros((0, 5), np.int64)

    # hanging-node finalisation (make_euler_mesh.py:787-950)
    insert = np.zeros((0, 2), np.int64)
    if pairs.shape[0]:
        f2c[pairs[:, 1:].ravel(), 1] = np.repeat(f2c[pairs[:, 0], 0], 4)
        keep_face = np.ones(nface, bool)
        keep_face[pairs[:, 0]] = False
        # coarse interface edges: their midpoint is an existing node (not an endpoint)
        elen = np.linalg.norm(nodes[e2n[:, 1]] - nodes[e2n[:, 0]], axis=1)
        eps = float(elen.min()) * 1e-3
        qn = np.round(nodes / eps).astype(np.int64)
        base = qn.min(axis=0)
        span = qn.max(axis=0) - base + 1
        if float(span[0]) * float(span[1]) * float(span[2]) >= 2.0 ** 62:
            raise ValueError("node lattice too large for packed keys")
        enc = lambda v: ((v[:, 0] - base[0]) * span[1] + (v[:, 1] - base[1])) * span[2] + (v[:, 2] - base[2])
        nkey = enc(qn)
        norder = np.argsort(nkey, kind="stable")
        nks = nkey[norder]
        probe = np.round((nodes[e2n[:, 0]] + 0.5 * (nodes[e2n[:, 1]] - nodes[e2n[:, 0]])) / eps).astype(np.int64)
        inside = np.all((probe >= base) & (probe < base + span), axis=1)
        pkey = enc(np.where(inside[:, None], probe, base))
        pos = np.minimum(np.searchsorted(nks, pkey), nks.shape[0] - 1)
        found = inside & (nks[pos] == pkey)
        mid = np.where(found, norder[pos], -1)
        coarse_e = found & (mid != e2n[:, 0]) & (mid != e2n[:, 1])
        ce = np.flatnonzero(coarse_e)
        m = mid[ce]
        n1, n2 = e2n[ce, 0], e2n[ce, 1]
        sk = _pack2(np.minimum(n1, m), np.maximum(n1, m))
        sk2 = _pack2(np.minimum(m, n2), np.maximum(m, n2))
        s1 = np.searchsorted(ukeys, sk)
        s2 = np.searchsorted(ukeys, sk2)
        s1c, s2c = np.minimum(s1, ukeys.shape[0] - 1), np.minimum(s2, ukeys.shape[0] - 1)
        ok = (ukeys[s1c] == sk) & (ukeys[s2c] == sk2)
        ce, s1, s2 = ce[ok], s1[ok], s2[ok]
        sub1 = np.full(e2n.shape[0], -1, np.int64)
        sub2 = np.full(e2n.shape[0], -1, np.int64)
        sub1[ce], sub2[ce] = s1, s2
        hit = keep_face[:, None] & (sub1[f2e] >= 0)
        hf, hs = np.nonzero(hit)                      # row-major: face order, then slot order
        ins_edge = sub2[f2e[hf, hs]]
        f2e[hf, hs] = sub1[f2e[hf, hs]]
        keep_edge = np.ones(e2n.shape[0], bool)
        keep_edge[ce] = False
        renum = np.full(e2n.shape[0], -1, np.int64)
        renum[keep_edge] = np.arange(int(keep_edge.sum()))
        f2e = renum[f2e]
        if (f2e[keep_face] < 0).any():
            raise ValueError("a kept face still references a removed coarse edge")
        e2n, bc_edge = e2n[keep_edge], bc_edge[keep_edge]
        frenum = np.full(nface, -1, np.int64)
        frenum[keep_face] = np.arange(int(keep_face.sum()))
        insert = np.column_stack([frenum[hf], renum[ins_edge]]).astype(np.int64)
        if (insert < 0).any():
            raise ValueError("link_fe_insert references a removed face or edge")
        f2n, f2c, f2e, bc_face = f2n[keep_face], f2c[keep_face], f2e[keep_face], bc_face[keep_face]

    nx, ny, nz = shape[0]
    if verbose:
        print(f"[mesh] {ncell:,} cells / {nxt:,} nodes / {f2n.shape[0]:,} faces / {e2n.shape[0]:,} edges")
    return CombinedMesh(
        nodes=nodes, cells=cells, n_layers=N, layer_cell_counts=tuple(layer_cells),
        layer_node_counts=tuple(nnode), refinement_ratio=2, boundary_faces=bfaces,
        boundary_fam=bfam, pos_ref=lo[0] + dx[0] / 2.0, dx_uniform=dx[0].copy(),
        index_cell=np.arange(nx * ny * nz, dtype=np.int64).reshape(nx, ny, nz),
        faces_to_nodes=np.ascontiguousarray(f2n), faces_to_cells=np.ascontiguousarray(f2c),
        edges_to_nodes=np.ascontiguousarray(e2n), faces_to_edges=np.ascontiguousarray(f2e),
        bc_node=bc_node, bc_edge=bc_edge, bc_face=bc_face, bc_cell=bc_cell,
        interface_pairs=pairs, link_fe_insert=insert,
        layer_spacing=tuple(float(d[0]) for d in dx), layer_shape=tuple(shape),
        layer_min=tuple(tuple(float(v) for v in l) for l in lo),
        layer_hole_lo=tuple(tuple(int(v) for v in hh) for hh in hole))


# Box lists: sphere_wake_mesh.py:46-59, sphere_quiescent_mesh.py:172-191, plus nested_box.
WAKE_BOUNDS_D = ((-1.0, 3.0, -1.0, 1.0, -1.0, 1.0), (-2.0, 6.0, -2.0, 2.0, -2.0, 2.0),
                 (-4.0, 12.0, -4.0, 4.0, -4.0, 4.0), (-8.0, 16.0, -8.0, 8.0, -8.0, 8.0))
CONCENTRIC_HALF_WIDTHS_D = (1.0, 2.0, 4.0, 8.0, 16.0)


def wake_boxes():
    return [BoxSpec.from_bounds(*tuple(DIAMETER * np.asarray(b))) for b in WAKE_BOUNDS_D]


def concentric_boxes():
    out = []
    for hw_d in CONCENTRIC_HALF_WIDTHS_D:
        hw = DIAMETER * hw_d
        out.append(BoxSpec.from_bounds(-hw, hw, -hw, hw, -hw, hw))
    return out


def nested_box_boxes(half_extents_d, levels):
    """Finest box = given half extents (diameters) about the origin; each coarser level doubles them."""
    out = []
    for k in range(int(levels)):
        hx, hy, hz = (DIAMETER * float(v) * (2 ** k) for v in half_extents_d)
        out.append(BoxSpec.from_bounds(-hx, hx, -hy, hy, -hz, hz))
    return out


def uniform_boxes(bounds):
    b = np.asarray(bounds, dtype=float)
    return [BoxSpec.from_bounds(float(b[0, 0]), float(b[0, 1]), float(b[1, 0]), float(b[1, 1]),
                                float(b[2, 0]), float(b[2, 1]))]


def closed_form_counts(shape, n_layers):
    """Cells/faces/edges of a nested mesh from its per-layer box shapes (each layer
    keeps its box minus the hole of the next-inner box)."""
    F = lambda n: (n[0] + 1) * n[1] * n[2] + n[0] * (n[1] + 1) * n[2] + n[0] * n[1] * (n[2] + 1)
    E = lambda n: n[0] * (n[1] + 1) * (n[2] + 1) + (n[0] + 1) * n[1] * (n[2] + 1) + (n[0] + 1) * (n[1] + 1) * n[2]
    C = lambda n: n[0] * n[1] * n[2]
    cells, faces, edges = C(shape[0]), F(shape[0]), E(shape[0])
    for k in range(1, n_layers):
        h = tuple(v // 2 for v in shape[k - 1])
        cells += C(shape[k]) - C(h)
        faces += F(shape[k]) - F(h)
        edges += E(shape[k]) - E(h)
    return cells, faces, edges
