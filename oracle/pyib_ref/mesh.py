"""Oracle restatement of the nested 2:1 Cartesian hex mesh builder.

TEST INFRASTRUCTURE ONLY (see package docstring).  Follows
``src/pyib/make_euler_mesh.py`` (``make_nested_mesh`` 330-463 and its helpers
466-711, 787-950, ``_build_box`` 1821-1830) and the box lists of
``src/pyib/solver/sphere_wake_mesh.py`` / ``sphere_quiescent_mesh.py``.  The
dictionary/loop based interface detection and hanging-node stitching are kept
in the reference's own element-by-element form, so this builder is slow on
large meshes but independent of the vectorised product builder.  It is pinned
by the sha256 layout fingerprints of the reference D32/D40 meshes.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# make_euler_mesh.py:211-227 -- hex corner order, local faces, local face edges
CORNERS = np.array([(1, 0, 1), (1, 0, 0), (0, 0, 1), (0, 0, 0),
                    (1, 1, 1), (1, 1, 0), (0, 1, 1), (0, 1, 0)], dtype=np.int64)
HEX_FACES = np.array([[3, 1, 0, 2], [7, 5, 4, 6], [3, 1, 5, 7],
                      [2, 0, 4, 6], [3, 7, 6, 2], [1, 5, 4, 0]], dtype=np.int64)
FACE_EDGES = np.array([[0, 1], [1, 2], [2, 3], [3, 0]], dtype=np.int64)


@dataclass(frozen=True)
class Box:
    """make_euler_mesh.py:134-161 (BoxSpec)."""
    center: tuple
    size: tuple

    @classmethod
    def bounds(cls, xl, xh, yl, yh, zl, zh):
        if not (xl < xh and yl < yh and zl < zh):
            raise ValueError("bounds need lo < hi on every axis")
        return cls(((xl + xh) / 2.0, (yl + yh) / 2.0, (zl + zh) / 2.0),
                   (xh - xl, yh - yl, zh - zl))

    @property
    def lo(self):
        return np.array(self.center, dtype=np.float64) - np.array(self.size, dtype=np.float64) / 2.0

    @property
    def hi(self):
        return np.array(self.center, dtype=np.float64) + np.array(self.size, dtype=np.float64) / 2.0


@dataclass(frozen=True)
class Mesh:
    """make_euler_mesh.py:164-195 (CombinedMesh); field names kept for the fingerprint."""
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


def _spec(h, boxes, r):
    """make_euler_mesh.py:234-290: per-layer spacing, cell counts, hole offsets."""
    h = np.atleast_1d(np.asarray(h, dtype=np.float64))
    if h.size == 1:
        h = np.array([float(h.flat[0])] * 3, dtype=np.float64)
    dx = [h * (r ** k) for k in range(len(boxes))]
    n_cells = []
    for k, b in enumerate(boxes):
        size = np.asarray(b.size, dtype=np.float64)
        n = np.round(size / dx[k]).astype(np.int64)
        if not np.allclose(size, n * dx[k], atol=1e-8):
            raise ValueError(f"layer {k} box size is not a multiple of its spacing")
        n_cells.append(tuple(int(v) for v in n))
    i_lo = [np.zeros(3, dtype=np.int64)]
    for k in range(1, len(boxes)):
        inner, outer = boxes[k - 1], boxes[k]
        if (inner.lo < outer.lo - 1e-9).any() or (inner.hi > outer.hi + 1e-9).any():
            raise ValueError(f"layer {k-1} box must lie inside layer {k} box")
        off = inner.lo - outer.lo
        oi = np.round(off / dx[k]).astype(np.int64)
        if not np.allclose(off, oi * dx[k], atol=1e-8):
            raise ValueError(f"layer {k-1} corners are not on layer {k}'s grid")
        i_lo.append(oi)
    return dx, n_cells, i_lo


def _cells_of_grid(ids, mask):
    """make_euler_mesh.py:466-478: 8 corner ids per live cell, C order."""
    nx, ny, nz = (s - 1 for s in ids.shape)
    corners = np.stack([ids[a:a + nx, b:b + ny, c:c + nz] for a, b, c in CORNERS], axis=-1)
    live = mask & (corners >= 0).all(axis=-1)
    return corners[live].astype(np.int64, copy=False)


def _outer_faces(g):
    """make_euler_mesh.py:481-531: outer-box boundary quads and families 1..6."""
    X, Y, Z = (s - 1 for s in g.shape)
    q = lambda a, b, c, d: np.stack([a, b, c, d], axis=-1).reshape(-1, 4)
    fams = [
        q(g[0, 0:Y, 1:Z + 1], g[0, 1:Y + 1, 1:Z + 1], g[0, 1:Y + 1, 0:Z], g[0, 0:Y, 0:Z]),
        q(g[X, 1:Y + 1, 1:Z + 1], g[X, 0:Y, 1:Z + 1], g[X, 0:Y, 0:Z], g[X, 1:Y + 1, 0:Z]),
        q(g[0:X, 0:Y, Z], g[1:X + 1, 0:Y, Z], g[1:X + 1, 1:Y + 1, Z], g[0:X, 1:Y + 1, Z]),
        q(g[0:X, 0:Y, 0], g[0:X, 1:Y + 1, 0], g[1:X + 1, 1:Y + 1, 0], g[1:X + 1, 0:Y, 0]),
        q(g[0:X, Y, 1:Z + 1], g[1:X + 1, Y, 1:Z + 1], g[1:X + 1, Y, 0:Z], g[0:X, Y, 0:Z]),
        q(g[1:X + 1, 0, 1:Z + 1], g[0:X, 0, 1:Z + 1], g[0:X, 0, 0:Z], g[1:X + 1, 0, 0:Z]),
    ]
    faces = np.concatenate(fams, axis=0).astype(np.int64, copy=False)
    fam = np.concatenate([np.full(f.shape[0], t, np.int64) for t, f in enumerate(fams, 1)])
    return faces, fam


def _connectivity(cells, bfaces, bfam, nnode):
    """make_euler_mesh.py:534-603: faces and edges numbered by lexsort of sorted node ids."""
    ncell = cells.shape[0]
    allf = cells[:, HEX_FACES].reshape(-1, 4)
    owner = np.repeat(np.arange(ncell, dtype=np.int64), 6)
    key = np.sort(allf, axis=1)
    order = np.lexsort(key.T[::-1])
    skey, sown, sorig = key[order], owner[order], allf[order]
    new = np.empty(skey.shape[0], bool)
    new[0] = True
    new[1:] = np.any(skey[1:] != skey[:-1], axis=1)
    gid = np.cumsum(new) - 1
    nface = int(gid[-1] + 1)
    cnt = np.bincount(gid, minlength=nface)
    if (cnt > 2).any():
        raise ValueError("non-manifold face")
    start = np.where(new)[0]
    f2n = sorig[start].copy()
    f2c = np.full((nface, 2), -1, np.int64)
    f2c[:, 0] = sown[start]
    two = cnt == 2
    f2c[two, 1] = sown[start[two] + 1]
    bc_face = np.zeros(nface, np.int64)
    ukey = skey[start]
    # look each boundary quad up among the unique keys (reference: packed-row searchsorted)
    lookup = {tuple(int(v) for v in k): i for i, k in enumerate(ukey)} if bfaces.shape[0] else {}
    for quad, fam in zip(np.sort(bfaces, axis=1), bfam):
        fid = lookup.get(tuple(int(v) for v in quad), -1)
        if fid < 0 or cnt[fid] != 1:
            raise ValueError("boundary quad is not a boundary face")
        bc_face[fid] = int(fam)
    fe = f2n[:, FACE_EDGES].reshape(-1, 2)
    ekey = np.sort(fe, axis=1)
    eorder = np.lexsort(ekey.T[::-1])
    sekey = ekey[eorder]
    enew = np.empty(sekey.shape[0], bool)
    enew[0] = True
    enew[1:] = np.any(sekey[1:] != sekey[:-1], axis=1)
    egid = np.cumsum(enew) - 1
    e2n = sekey[np.where(enew)[0]].copy()
    eid = np.empty(fe.shape[0], np.int64)
    eid[eorder] = egid
    f2e = eid.reshape(nface, 4)
    bc_node = np.zeros(nnode, np.int64)
    bc_edge = np.zeros(e2n.shape[0], np.int64)
    bc_cell = np.zeros(ncell, np.int64)
    for f in np.where(bc_face != 0)[0]:
        fam = bc_face[f]
        for n in f2n[f]:
            bc_node[n] = max(bc_node[n], fam)
        for e in f2e[f]:
            bc_edge[e] = max(bc_edge[e], fam)
    return f2n, f2c, e2n, f2e, bc_node, bc_edge, bc_face, bc_cell


def _interface_pairs(cells, layer_cells, n_cells, i_lo, r, f2n):
    """make_euler_mesh.py:606-709: coarse interface face -> its r*r fine sub-faces."""
    key_to_face = {tuple(int(v) for v in k): i for i, k in enumerate(np.sort(f2n, axis=1))}
    sides = (("x-", 0, False, 4, 5), ("x+", 0, True, 5, 4), ("y-", 1, False, 0, 1),
             ("y+", 1, True, 1, 0), ("z-", 2, False, 2, 3), ("z+", 2, True, 3, 2))
    pairs = []
    for k in range(len(layer_cells) - 1):
        a_lo = i_lo[k + 1]
        nk = np.asarray(n_cells[k], np.int64)
        a_hi = a_lo + nk // r
        # global ids of the live coarse (layer k+1) and fine (layer k) cells
        ncp = n_cells[k + 1]
        alive_c = np.ones(ncp, bool)
        alive_c[a_lo[0]:a_hi[0], a_lo[1]:a_hi[1], a_lo[2]:a_hi[2]] = False
        cid_c = np.full(ncp, -1, np.int64)
        off_c = sum(layer_cells[:k + 1])
        cid_c[alive_c] = np.arange(off_c, off_c + int(alive_c.sum()))
        alive_f = np.ones(n_cells[k], bool)
        if k > 0:
            h0 = i_lo[k]
            hn = np.asarray(n_cells[k - 1], np.int64) // r
            alive_f[h0[0]:h0[0] + hn[0], h0[1]:h0[1] + hn[1], h0[2]:h0[2] + hn[2]] = False
        cid_f = np.full(n_cells[k], -1, np.int64)
        off_f = sum(layer_cells[:k])
        cid_f[alive_f] = np.arange(off_f, off_f + int(alive_f.sum()))
        for _name, axis, is_max, fine_lf, coarse_lf in sides:
            o0, o1 = [a for a in (0, 1, 2) if a != axis]
            cax = a_hi[axis] if is_max else a_lo[axis] - 1
            fax = nk[axis] - 1 if is_max else 0
            for c0 in range(a_lo[o0], a_hi[o0]):
                for c1 in range(a_lo[o1], a_hi[o1]):
                    ci = [0, 0, 0]
                    ci[axis], ci[o0], ci[o1] = cax, c0, c1
                    coarse = cid_c[tuple(ci)]
                    if coarse < 0:
                        continue
                    cf = key_to_face.get(tuple(int(v) for v in np.sort(cells[coarse][HEX_FACES[coarse_lf]])), -1)
                    if cf < 0:
                        continue
                    fine = []
                    for d0 in range(r):
                        for d1 in range(r):
                            fi = [0, 0, 0]
                            fi[axis] = fax
                            fi[o0] = (c0 - a_lo[o0]) * r + d0
                            fi[o1] = (c1 - a_lo[o1]) * r + d1
                            fc = cid_f[tuple(fi)]
                            ff = -1 if fc < 0 else key_to_face.get(
                                tuple(int(v) for v in np.sort(cells[fc][HEX_FACES[fine_lf]])), -1)
                            if ff < 0:
                                fine = None
                                break
                            fine.append(ff)
                        if fine is None:
                            break
                    if fine is not None and len(fine) == r * r:
                        pairs.append([cf] + fine)
    if not pairs:
        return np.zeros((0, 1 + r * r), np.int64)
    return np.array(pairs, np.int64)


def _stitch(f2n, f2c, f2e, e2n, bc_face, bc_edge, pairs, nodes, r):
    """make_euler_mesh.py:787-950: face aliasing, coarse-edge collapse, link_fe_insert."""
    keep_face = np.ones(f2n.shape[0], bool)
    f2c_new = f2c.copy()
    if pairs.shape[0]:
        coarse_cell = f2c[pairs[:, 0], 0]
        for j in range(1, pairs.shape[1]):
            if (f2c_new[pairs[:, j], 1] != -1).any():
                raise ValueError("fine interface face already has a second cell")
            f2c_new[pairs[:, j], 1] = coarse_cell
        keep_face[pairs[:, 0]] = False
    ne = e2n.shape[0]
    f2e_new = f2e.copy()
    insert = np.zeros((0, 2), np.int64)
    e2n_new, bc_edge_new = e2n, bc_edge
    if r >= 2 and pairs.shape[0]:
        eps = float(np.linalg.norm(nodes[e2n[:, 1]] - nodes[e2n[:, 0]], axis=1).min()) * 1e-3
        node_at = {tuple(k): i for i, k in enumerate(map(tuple, np.round(nodes / eps).astype(np.int64)))}
        probe = np.round((nodes[e2n[:, 0]] + (1.0 / r) * (nodes[e2n[:, 1]] - nodes[e2n[:, 0]])) / eps).astype(np.int64)
        coarse_edges = []
        for e in range(ne):
            hn = node_at.get(tuple(probe[e]))
            if hn is not None and hn != int(e2n[e, 0]) and hn != int(e2n[e, 1]):
                coarse_edges.append(e)
        if coarse_edges:
            edge_at = {(min(int(a), int(b)), max(int(a), int(b))): i for i, (a, b) in enumerate(e2n)}
            subs = {}
            for ce in coarse_edges:
                n1, n2 = int(e2n[ce, 0]), int(e2n[ce, 1])
                pts = [n1]
                for i in range(1, r):
                    key = tuple(np.round((nodes[n1] + (i / r) * (nodes[n2] - nodes[n1])) / eps).astype(np.int64))
                    if key not in node_at:
                        pts = None
                        break
                    pts.append(node_at[key])
                if pts is None:
                    continue
                pts.append(n2)
                chain = []
                for j in range(r):
                    a, b = pts[j], pts[j + 1]
                    eid = edge_at.get((min(a, b), max(a, b)))
                    if eid is None:
                        chain = None
                        break
                    chain.append(eid)
                if chain is not None:
                    subs[int(ce)] = chain
            rows = []
            for f in range(f2e.shape[0]):
                if not keep_face[f]:
                    continue
                for slot in range(f2e.shape[1]):
                    ce = int(f2e_new[f, slot])
                    if ce in subs:
                        f2e_new[f, slot] = subs[ce][0]
                        rows.extend((f, fe) for fe in subs[ce][1:])
            keep_edge = np.ones(ne, bool)
            keep_edge[list(subs.keys())] = False
            renum = np.full(ne, -1, np.int64)
            renum[keep_edge] = np.arange(int(keep_edge.sum()))
            f2e_new = renum[f2e_new]
            if (f2e_new[keep_face] < 0).any():
                raise ValueError("kept face references a collapsed edge")
            e2n_new, bc_edge_new = e2n[keep_edge], bc_edge[keep_edge]
            insert = np.array(rows, np.int64) if rows else np.zeros((0, 2), np.int64)
            if insert.size:
                insert[:, 1] = renum[insert[:, 1]]
    frenum = np.full(f2n.shape[0], -1, np.int64)
    frenum[keep_face] = np.arange(int(keep_face.sum()))
    if insert.size:
        insert[:, 0] = frenum[insert[:, 0]]
        if (insert < 0).any():
            raise ValueError("link_fe_insert references a removed face")
    return (f2n[keep_face], f2c_new[keep_face], f2e_new[keep_face], bc_face[keep_face],
            e2n_new, bc_edge_new, insert)


def make_nested_mesh(h, boxes, refinement_ratio=2):
    """make_euler_mesh.py:330-463."""
    dx, n_cells, i_lo = _spec(h, boxes, refinement_ratio)
    r = int(refinement_ratio)
    N = len(boxes)
    nshape = [(n[0] + 1, n[1] + 1, n[2] + 1) for n in n_cells]
    lo = [b.lo for b in boxes]
    ids = [None] * N
    nnodes = [0] * N
    nxt = 0

    def hole_interior(k):
        # nodes strictly inside the next-inner box are not created on layer k
        m = np.zeros(nshape[k], bool)
        if k > 0:
            hn = np.round(np.asarray(boxes[k - 1].size) / dx[k]).astype(np.int64)
            a = i_lo[k]
            m[a[0] + 1:a[0] + hn[0], a[1] + 1:a[1] + hn[1], a[2] + 1:a[2] + hn[2]] = True
        return m

    k = N - 1
    ids[k] = np.full(nshape[k], -1, np.int64)
    live = ~hole_interior(k)
    n = int(live.sum())
    ids[k][live] = np.arange(nxt, nxt + n)
    nnodes[k] = n
    nxt += n
    for k in range(N - 2, -1, -1):
        ids[k] = np.full(nshape[k], -1, np.int64)
        a, b, c = np.indices(nshape[k])
        same = (a % r == 0) & (b % r == 0) & (c % r == 0)
        o = i_lo[k + 1]
        ids[k][same] = ids[k + 1][(o[0] + a // r)[same], (o[1] + b // r)[same], (o[2] + c // r)[same]]
        need = (ids[k] < 0) & ~hole_interior(k)
        n = int(need.sum())
        ids[k][need] = np.arange(nxt, nxt + n)
        nnodes[k] = n
        nxt += n
    nodes = np.zeros((nxt, 3), np.float64)
    for k in range(N):
        a, b, c = np.where(ids[k] >= 0)
        nodes[ids[k][a, b, c]] = np.column_stack([lo[k][0] + a * dx[k][0], lo[k][1] + b * dx[k][1],
                                                  lo[k][2] + c * dx[k][2]])
    per_layer = []
    for k in range(N):
        mask = np.ones(n_cells[k], bool)
        if k > 0:
            hn = np.round(np.asarray(boxes[k - 1].size) / dx[k]).astype(np.int64)
            a = i_lo[k]
            mask[a[0]:a[0] + hn[0], a[1]:a[1] + hn[1], a[2]:a[2] + hn[2]] = False
        per_layer.append(_cells_of_grid(ids[k], mask))
    layer_cells = [int(c.shape[0]) for c in per_layer]
    cells = np.concatenate(per_layer, axis=0)
    bfaces, bfam = _outer_faces(ids[N - 1])
    pos_ref = lo[0] + dx[0] / 2.0
    nx, ny, nz = n_cells[0]
    index_cell = np.arange(nx * ny * nz, dtype=np.int64).reshape(nx, ny, nz)
    f2n, f2c, e2n, f2e, bc_node, bc_edge, bc_face, bc_cell = _connectivity(cells, bfaces, bfam, nxt)
    if r > 1 and N > 1:
        pairs = _interface_pairs(cells, layer_cells, n_cells, i_lo, r, f2n)
    else:
        pairs = np.zeros((0, 1 + r * r), np.int64)
    if r > 1 and N > 1 and pairs.shape[0]:
        f2n, f2c, f2e, bc_face, e2n, bc_edge, insert = _stitch(
            f2n, f2c, f2e, e2n, bc_face, bc_edge, pairs, nodes, r)
    else:
        insert = np.zeros((0, 2), np.int64)
    return Mesh(nodes=nodes, cells=cells, n_layers=N, layer_cell_counts=tuple(layer_cells),
                layer_node_counts=tuple(nnodes), refinement_ratio=r, boundary_faces=bfaces,
                boundary_fam=bfam, pos_ref=pos_ref, dx_uniform=dx[0].copy(), index_cell=index_cell,
                faces_to_nodes=f2n, faces_to_cells=f2c, edges_to_nodes=e2n, faces_to_edges=f2e,
                bc_node=bc_node, bc_edge=bc_edge, bc_face=bc_face, bc_cell=bc_cell,
                interface_pairs=pairs, link_fe_insert=insert)


# sphere_wake_mesh.py:42-51 and sphere_quiescent_mesh.py:39-41
DIAMETER = 0.5
WAKE_BOUNDS_D = ((-1.0, 3.0, -1.0, 1.0, -1.0, 1.0), (-2.0, 6.0, -2.0, 2.0, -2.0, 2.0),
                 (-4.0, 12.0, -4.0, 4.0, -4.0, 4.0), (-8.0, 16.0, -8.0, 8.0, -8.0, 8.0))
CONCENTRIC_HALF_WIDTHS_D = (1.0, 2.0, 4.0, 8.0, 16.0)


def wake_boxes():
    """sphere_wake_mesh.py:54-59."""
    return [Box.bounds(*tuple(DIAMETER * np.asarray(b))) for b in WAKE_BOUNDS_D]


def concentric_boxes():
    """sphere_quiescent_mesh.py:44-60."""
    out = []
    for hw_d in CONCENTRIC_HALF_WIDTHS_D:
        hw = DIAMETER * hw_d
        out.append(Box.bounds(-hw, hw, -hw, hw, -hw, hw))
    return out


def nested_box_boxes(half_extents_d, levels):
    """Additive mesh kind ``nested_box`` (no reference counterpart): the finest box
    has the given half extents (in diameters) about the origin and every coarser
    level doubles them."""
    out = []
    for k in range(levels):
        hx, hy, hz = (DIAMETER * float(v) * (2 ** k) for v in half_extents_d)
        out.append(Box.bounds(-hx, hx, -hy, hy, -hz, hz))
    return out


def uniform_box(bounds):
    """ib_case.py:606-611 + 644-648: one-layer box from [[lo,hi]]*3."""
    b = np.asarray(bounds, dtype=float)
    return [Box.bounds(float(b[0, 0]), float(b[0, 1]), float(b[1, 0]), float(b[1, 1]),
                       float(b[2, 0]), float(b[2, 1]))]
