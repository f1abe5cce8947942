"""Oracle restatement of the mesh geometry and incidence operators (IBCore).

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_core_np.py`` in float64
(``IBCore.__init__`` 68-171, operators 174-228).  Every reduction is a NumPy
``bincount``/``add.at`` in link order, as in the reference CPU backend.
"""
from __future__ import annotations

import numpy as np


def _avg_to_cells(F2C, m1, ncell, fvals):
    """Cell average of face quantities over both owners (ib_core_np.py:50-53, 78-82)."""
    s = np.zeros((ncell, 3))
    c = np.zeros(ncell)
    np.add.at(s, F2C[:, 0], fvals)
    np.add.at(c, F2C[:, 0], 1.0)
    np.add.at(s, F2C[m1, 1], fvals[m1])
    np.add.at(c, F2C[m1, 1], 1.0)
    return s / c[:, None]


class Core:
    """ib_core_np.py:24-126 (IBCore) in float64."""

    def __init__(self, mesh, ND=3, geometry=None):
        self.ND = ND
        if geometry is not None:
            # Geometry already computed by an equivalent builder (bitwise-equal, checked by the
            # tests): reuse it so that a timing run does not pay the set-up twice.
            for name in ("F2C", "ncell", "nface", "nedge", "m1", "lf_face", "lf_edge", "epos", "te",
                         "fe_sign_link", "nf", "Af", "fpos", "cpos", "rf1", "rf2", "rf", "vol_ci", "AoL"):
                setattr(self, name, getattr(geometry, name))
            self.E2N = mesh.edges_to_nodes
            return
        nodes = mesh.nodes
        self.E2N = E2N = mesh.edges_to_nodes
        self.F2C = F2C = mesh.faces_to_cells
        F2E = mesh.faces_to_edges
        self.ncell = ncell = mesh.cells.shape[0]
        self.nface = nface = F2C.shape[0]
        self.nedge = E2N.shape[0]
        self.m1 = m1 = F2C[:, 1] >= 0
        # ib_core_np.py:39-41 -- flat face/edge links: 4 per face + link_fe_insert
        ins = np.asarray(getattr(mesh, "link_fe_insert", np.zeros((0, 2), np.int64)))
        self.lf_face = lf_face = np.concatenate([np.repeat(np.arange(nface), 4), ins[:, 0]])
        self.lf_edge = lf_edge = np.concatenate([F2E.ravel(), ins[:, 1]])
        # ib_core_np.py:44-53 -- edge midpoint/tangent, face and cell centres
        self.epos = epos = nodes[E2N].mean(axis=1)
        self.te = te = nodes[E2N[:, 1]] - nodes[E2N[:, 0]]
        fs = np.zeros((nface, 3))
        fc = np.zeros(nface)
        np.add.at(fs, lf_face, epos[lf_edge])
        np.add.at(fc, lf_face, 1.0)
        fpos = fs / fc[:, None]
        cpos = _avg_to_cells(F2C, m1, ncell, fpos)
        # ib_core_np.py:55-62 -- pre-correction rf
        rf = np.zeros((nface, 3))
        rf[m1] = cpos[F2C[m1, 1]] - cpos[F2C[m1, 0]]
        rf[~m1] = fpos[~m1] - cpos[F2C[~m1, 0]]
        # ib_core_np.py:64-72 -- per-link circulation sign and face area vector
        avec = np.cross(epos[lf_edge] - fpos[lf_face], te[lf_edge])
        sgn = np.sign((avec * rf[lf_face]).sum(axis=1))
        sgn[sgn == 0] = 1.0
        self.fe_sign_link = sgn
        nf = np.zeros((nface, 3))
        np.add.at(nf, lf_face, avec * sgn[:, None])
        nf *= 1.0 / (ND - 1)
        self.nf = nf
        self.Af = Af = np.linalg.norm(nf, axis=1)
        # ib_core_np.py:74-102 -- Modify_geom (2:1 hanging-node centre corrections)
        tol = 1e-6 * Af
        rfe = epos[lf_edge] - fpos[lf_face]
        hitf = np.abs((rfe * rfe).sum(1) - 0.36 * Af[lf_face]) < tol[lf_face]
        np.add.at(fpos, lf_face[hitf], rfe[hitf] / 6.0)
        cpos = _avg_to_cells(F2C, m1, ncell, fpos)
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
        # ib_core_np.py:104-111 -- inverse cell volume and diffusion face factor
        vraw = np.zeros(ncell)
        np.add.at(vraw, F2C[:, 0], (nf * rf1).sum(axis=1))
        np.add.at(vraw, F2C[m1, 1], -(nf[m1] * rf2[m1]).sum(axis=1))
        self.vol_ci = np.where(np.abs(vraw) > 1e-14, ND / vraw, 0.0)
        self.AoL = (rf * nf).sum(axis=1) / (rf * rf).sum(axis=1)
        # ib_core_np.py:113-121 -- gradient-reconstruction matrix (self-test only)
        contrib = 0.5 * nf[:, :, None] * rf[:, None, :]
        matr = np.zeros((ncell, 3, 3))
        np.add.at(matr, F2C[:, 0], contrib)
        np.add.at(matr, F2C[m1, 1], contrib[m1])
        matr[self.vol_ci <= 0] = np.eye(3)
        self.matr = matr

    # ib_core_np.py:129-183 -------------------------------------------------
    def rot(self, face):
        return np.bincount(self.lf_edge, self.fe_sign_link * face[self.lf_face], minlength=self.nedge)

    def curl(self, edge):
        return np.bincount(self.lf_face, self.fe_sign_link * edge[self.lf_edge], minlength=self.nface)

    def cdiv(self, face):
        m1 = self.m1
        return (np.bincount(self.F2C[:, 0], face, minlength=self.ncell)
                - np.bincount(self.F2C[m1, 1], face[m1], minlength=self.ncell))

    def integr_c2f(self, cellvec):
        out = (cellvec[self.F2C[:, 0]] * self.rf1).sum(axis=1)
        m1 = self.m1
        out[m1] -= (cellvec[self.F2C[m1, 1]] * self.rf2[m1]).sum(axis=1)
        return out

    def interp_f2c(self, face):
        m1 = self.m1
        c0, c1 = self.F2C[:, 0], self.F2C[m1, 1]
        out = np.empty((self.ncell, 3))
        for d in range(3):
            out[:, d] = (np.bincount(c0, face * self.rf1[:, d], minlength=self.ncell)
                         - np.bincount(c1, face[m1] * self.rf2[m1, d], minlength=self.ncell))
        return out * self.vol_ci[:, None]

    def grad_recon(self, face_jump):
        m1 = self.m1
        w = 0.5 * face_jump
        rhs = np.empty((self.ncell, 3))
        for d in range(3):
            rhs[:, d] = (np.bincount(self.F2C[:, 0], w * self.nf[:, d], minlength=self.ncell)
                         + np.bincount(self.F2C[m1, 1], w[m1] * self.nf[m1, d], minlength=self.ncell))
        return np.linalg.solve(self.matr, rhs[..., None])[..., 0]

    def cgrad_cellscalar(self, cell):
        out = np.empty(self.nface)
        m1 = self.m1
        out[m1] = cell[self.F2C[m1, 1]] - cell[self.F2C[m1, 0]]
        out[~m1] = -cell[self.F2C[~m1, 0]]
        return out

    def flow_mat(self, pp, alpd=0.0, visc=0.0, rou=1.0, dt=0.0):
        """ib_core_np.py:168-175; the viscous factor is the literal 0.5."""
        vcq = self.interp_f2c(self.curl(pp))
        if alpd != 0.0:
            for k in range(self.ND):
                qf = self.AoL * visc * self.cgrad_cellscalar(vcq[:, k])
                vcq[:, k] = rou * vcq[:, k] - 0.5 * dt * self.vol_ci * self.cdiv(qf)
        return self.rot(self.integr_c2f(vcq))
