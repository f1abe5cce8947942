"""Host-side mesh geometry of the stream-function operators (setup, float64).

Computes the same quantities as the reference ``IBCore.__init__``
(``src/pyib/solver/ib_core_np.py:24-126``): the flat face/edge link list
(4 links per face + ``link_fe_insert``), edge midpoints/tangents, face and cell
centres, the per-link circulation sign, face area vectors, the 2:1
"Modify_geom" centre corrections, ``rf/rf1/rf2``, ``vol_ci`` and ``AoL``.
The work runs in C++ (``pyib_geometry`` in ``csrc/geometry.cpp``). Its sums keep the
reference's link order, so every value is bitwise the NumPy statement's. Everything here
runs once per mesh; the device kernels read the derived coefficient tables built in
``tables.py``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Geometry:
    def __init__(self, mesh, ND=3):
        if ND != 3:
            raise ValueError("the geometry is three-dimensional")
        self.ND = ND
        nodes = _c(mesh.nodes, np.float64)
        e2n = _c(mesh.edges_to_nodes, np.int64)
        self.F2C = f2c = _c(mesh.faces_to_cells, np.int64)
        f2e = _c(mesh.faces_to_edges, np.int64)
        ins = _c(np.asarray(mesh.link_fe_insert).reshape(-1, 2), np.int64)
        self.ncell = ncell = int(mesh.cells.shape[0])
        self.nface = nface = int(f2c.shape[0])
        self.nedge = nedge = int(e2n.shape[0])
        self.m1 = f2c[:, 1] >= 0
        nlink = 4 * nface + ins.shape[0]
        self.lf_face = np.empty(nlink, np.int64)
        self.lf_edge = np.empty(nlink, np.int64)
        self.epos, self.te = np.empty((nedge, 3)), np.empty((nedge, 3))
        self.fpos, self.cpos = np.empty((nface, 3)), np.empty((ncell, 3))
        self.fe_sign_link = np.empty(nlink)
        self.nf, self.Af = np.empty((nface, 3)), np.empty(nface)
        self.rf1, self.rf2, self.rf = np.empty((nface, 3)), np.empty((nface, 3)), np.empty((nface, 3))
        self.vol_ci, self.AoL = np.empty(ncell), np.empty(nface)
        N.call("pyib_geometry", C.c_int64(nodes.shape[0]), _p(nodes), C.c_int64(nedge), _p(e2n), C.c_int64(nface),
               _p(f2c), _p(f2e), C.c_int64(ins.shape[0]), _p(ins), C.c_int64(ncell), _p(self.lf_face),
               _p(self.lf_edge), _p(self.epos), _p(self.te), _p(self.fpos), _p(self.cpos), _p(self.fe_sign_link),
               _p(self.nf), _p(self.Af), _p(self.rf1), _p(self.rf2), _p(self.rf), _p(self.vol_ci), _p(self.AoL))


def diag_preconditioner(geo, dir_mask):
    """Inverse diagonal Jacobi preconditioner (``ib_driver_np.py:78-88``)."""
    qf = (geo.rf * geo.nf).sum(axis=1) / (geo.Af * geo.Af)
    prec = np.bincount(geo.lf_edge, np.abs(geo.fe_sign_link) * qf[geo.lf_face], minlength=geo.nedge)
    prec[dir_mask] = 0.0
    inv = np.zeros(geo.nedge)
    nz = prec != 0.0
    inv[nz] = 1.0 / prec[nz]
    return inv
