"""Nested 2:1 Cartesian hex mesh in PyIB's global numbering (host setup).

Builds the same arrays as the reference ``make_nested_mesh``
(``src/pyib/make_euler_mesh.py:330-463``): node ids assigned outermost layer
first, cells layer by layer in C order, faces and edges numbered by the
lexicographic order of their sorted node ids, then the hanging-node
finalisation of ``make_euler_mesh.py:787-950`` (fine interface faces take the
coarse cell as second owner, coarse interface faces are removed, each coarse
interface edge is collapsed onto its two fine halves with the second half
recorded in ``link_fe_insert``).  The arrays are built by the native builder
``pyib_mesh_build`` (``csrc/mesh.cpp``), which enumerates each face and edge
once from its lattice position and numbers them by counting sorts, so the
40M-edge meshes of the 3-level D32/D40 oscillating-sphere cases build in
seconds.  The ordering is bit-identical to the reference, which the sha256
layout fingerprint of the reference D32/D40 meshes pins.

Only refinement ratio 2 is supported: the geometry correction of the solver
(``ib_core_np.py:74-102``) is written for ratio 2.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

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


def make_nested_mesh(inner_spacing, boxes, refinement_ratio=2, *, verbose=False) -> CombinedMesh:
    """The nested mesh of ``boxes`` (finest first) with finest spacing ``inner_spacing``, built by
    the native builder ``pyib_mesh_build`` (``csrc/mesh.cpp``)."""
    import ctypes as C

    from . import _native as N

    if int(refinement_ratio) != 2:
        raise ValueError("only refinement_ratio=2 is supported")
    dx, shape, hole = _layers(inner_spacing, boxes)
    N_ = len(boxes)
    lo = [b.min for b in boxes]
    shp = np.ascontiguousarray(np.asarray(shape, np.int64).reshape(N_, 3))
    lo_a = np.ascontiguousarray(np.asarray(lo, np.float64).reshape(N_, 3))
    dx_a = np.ascontiguousarray(np.asarray(dx, np.float64).reshape(N_, 3))
    hole_a = np.ascontiguousarray(np.asarray(hole, np.int64).reshape(N_, 3))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    handle = C.c_void_p()
    counts = np.zeros(8, np.int64)
    N.call("pyib_mesh_build", C.c_int32(N_), p(shp), p(lo_a), p(dx_a), p(hole_a), C.byref(handle), p(counts))
    try:
        nnode, ncell, nbf, nface, nedge, npair, nins = (int(v) for v in counts[:7])
        nodes = np.empty((nnode, 3), np.float64)
        cells = np.empty((ncell, 8), np.int64)
        bfaces, bfam = np.empty((nbf, 4), np.int64), np.empty(nbf, np.int64)
        f2n, f2c, f2e = np.empty((nface, 4), np.int64), np.empty((nface, 2), np.int64), np.empty((nface, 4), np.int64)
        e2n = np.empty((nedge, 2), np.int64)
        bc_node, bc_edge, bc_face = np.empty(nnode, np.int64), np.empty(nedge, np.int64), np.empty(nface, np.int64)
        pairs, insert = np.empty((npair, 5), np.int64), np.empty((nins, 2), np.int64)
        lc, ln = np.empty(N_, np.int64), np.empty(N_, np.int64)
        N.call("pyib_mesh_fetch", handle, p(nodes), p(cells), p(bfaces), p(bfam), p(f2n), p(f2c), p(e2n), p(f2e),
               p(bc_node), p(bc_edge), p(bc_face), p(pairs), p(insert), p(lc), p(ln))
    finally:
        N.call("pyib_mesh_free", handle)
    nx, ny, nz = shape[0]
    if verbose:
        print(f"[mesh] {ncell:,} cells / {nnode:,} nodes / {nface:,} faces / {nedge:,} edges")
    return CombinedMesh(
        nodes=nodes, cells=cells, n_layers=N_, layer_cell_counts=tuple(int(v) for v in lc),
        layer_node_counts=tuple(int(v) for v in ln), refinement_ratio=2, boundary_faces=bfaces,
        boundary_fam=bfam, pos_ref=lo[0] + dx[0] / 2.0, dx_uniform=dx[0].copy(),
        index_cell=np.arange(nx * ny * nz, dtype=np.int64).reshape(nx, ny, nz),
        faces_to_nodes=f2n, faces_to_cells=f2c, edges_to_nodes=e2n, faces_to_edges=f2e,
        bc_node=bc_node, bc_edge=bc_edge, bc_face=bc_face, bc_cell=np.zeros(ncell, np.int64),
        interface_pairs=pairs, link_fe_insert=insert,
        layer_spacing=tuple(float(d[0]) for d in dx), layer_shape=tuple(shape),
        layer_min=tuple(tuple(float(v) for v in l) for l in lo),
        layer_hole_lo=tuple(tuple(int(v) for v in hh) for hh in hole))


# Box lists: sphere_wake_mesh.py:46-59, sphere_quiescent_mesh.py:41-60, plus nested_box.
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
