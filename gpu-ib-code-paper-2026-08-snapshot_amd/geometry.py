"""Host-side mesh geometry of the stream-function operators (setup, float64).

Computes the same quantities as the reference ``IBCore.__init__``
(``src/pyib/solver/ib_core_np.py:68-166``): the flat face/edge link list
(4 links per face + ``link_fe_insert``), edge midpoints/tangents, face and cell
centres, the per-link circulation sign, face area vectors, the 2:1
"Modify_geom" centre corrections, ``rf/rf1/rf2``, ``vol_ci`` and ``AoL``.
Reductions are ``np.bincount`` in link order, which sums in the same order as
the reference's ``np.add.at``.  Everything here runs once per mesh; the device
kernels read the derived coefficient tables built in ``tables.py``.
"""
from __future__ import annotations

import numpy as np


def _bc3(idx, vals, n):
    return np.stack([np.bincount(idx, vals[:, d], minlength=n) for d in range(3)], axis=1)


class Geometry:
    def __init__(self, mesh, ND=3):
        self.ND = ND
        nodes = mesh.nodes
        E2N = mesh.edges_to_nodes
        self.F2C = F2C = mesh.faces_to_cells
        F2E = mesh.faces_to_edges
        self.ncell = ncell = int(mesh.cells.shape[0])
        self.nface = nface = int(F2C.shape[0])
        self.nedge = int(E2N.shape[0])
        self.m1 = m1 = F2C[:, 1] >= 0
        ins = np.asarray(mesh.link_fe_insert)
        self.lf_face = lf_face = np.concatenate([np.repeat(np.arange(nface), 4), ins[:, 0]])
        self.lf_edge = lf_edge = np.concatenate([F2E.ravel(), ins[:, 1]])
        self.epos = epos = 0.5 * (nodes[E2N[:, 0]] + nodes[E2N[:, 1]])
        self.te = te = nodes[E2N[:, 1]] - nodes[E2N[:, 0]]
        fcnt = np.bincount(lf_face, minlength=nface).astype(np.float64)
        fpos = _bc3(lf_face, epos[lf_edge], nface) / fcnt[:, None]
        c1, c2 = F2C[:, 0], F2C[m1, 1]

        cboth = np.concatenate([c1, c2])        # c1 owners first, then c2: the reference's add order
        ccnt = np.bincount(cboth, minlength=ncell).astype(np.float64)

        def cells_avg(fp):
            return _bc3(cboth, np.concatenate([fp, fp[m1]]), ncell) / ccnt[:, None]

        cpos = cells_avg(fpos)
        rf = np.empty((nface, 3))
        rf[m1] = cpos[F2C[m1, 1]] - cpos[F2C[m1, 0]]
        rf[~m1] = fpos[~m1] - cpos[F2C[~m1, 0]]
        avec = np.cross(epos[lf_edge] - fpos[lf_face], te[lf_edge])
        sgn = np.sign((avec * rf[lf_face]).sum(axis=1))
        sgn[sgn == 0] = 1.0
        self.fe_sign_link = sgn
        nf = _bc3(lf_face, avec * sgn[:, None], nface) * (1.0 / (ND - 1))
        del avec
        self.nf = nf
        self.Af = Af = np.linalg.norm(nf, axis=1)
        # 2:1 interface centre corrections (ib_core_np.py:123-147)
        tol = 1e-6 * Af
        rfe = epos[lf_edge] - fpos[lf_face]
        hitf = np.abs((rfe * rfe).sum(1) - 0.36 * Af[lf_face]) < tol[lf_face]
        np.add.at(fpos, lf_face[hitf], rfe[hitf] / 6.0)
        del rfe
        cpos = cells_avg(fpos)
        rf1 = fpos - cpos[F2C[:, 0]]
        rf2 = np.zeros((nface, 3))
        rf2[m1] = fpos[m1] - cpos[F2C[m1, 1]]
        a49 = (4.0 / 9.0) * Af
        h1 = np.abs((rf1 * rf1).sum(1) - a49) < tol
        np.add.at(cpos, F2C[h1, 0], rf1[h1] / 4.0)
        h2 = (~h1) & m1 & (np.abs((rf2 * rf2).sum(1) - a49) < tol)
        np.add.at(cpos, F2C[h2, 1], rf2[h2] / 4.0)
        self.fpos, self.cpos = fpos, cpos
        self.rf1 = rf1 = fpos - cpos[F2C[:, 0]]
        self.rf2 = rf2 = np.zeros((nface, 3))
        rf2[m1] = fpos[m1] - cpos[F2C[m1, 1]]
        self.rf = rf = rf1 - rf2
        vraw = np.bincount(cboth, np.concatenate([(nf * rf1).sum(axis=1), -(nf[m1] * rf2[m1]).sum(axis=1)]),
                           minlength=ncell)
        self.vol_ci = np.where(np.abs(vraw) > 1e-14, ND / vraw, 0.0)
        self.AoL = (rf * nf).sum(axis=1) / (rf * rf).sum(axis=1)


def diag_preconditioner(geo, dir_mask):
    """Inverse diagonal Jacobi preconditioner (``ib_driver_np.py:78-88``)."""
    qf = (geo.rf * geo.nf).sum(axis=1) / (geo.Af * geo.Af)
    prec = np.bincount(geo.lf_edge, np.abs(geo.fe_sign_link) * qf[geo.lf_face], minlength=geo.nedge)
    prec[dir_mask] = 0.0
    inv = np.zeros(geo.nedge)
    nz = prec != 0.0
    inv[nz] = 1.0 / prec[nz]
    return inv
