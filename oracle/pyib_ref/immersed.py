"""Oracle restatement of Roma-Peskin 3-point spreading and interpolation.

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_immersed.py``:
``delta_h3`` 12-19, stencil build 63-81, interp 83-93, spread 95-111,
``direct_force`` 113-120, implicit forcing ``prepare_implicit`` / ``implicit_force``
122-174.  Spreading is a ``bincount`` over the flattened (marker, tap) list, i.e. a
sequential sum in marker order.
"""
from __future__ import annotations

import numpy as np


def delta_h3(r, h):
    """ib_immersed.py:12-19."""
    a = np.abs(r / h)
    out = np.zeros_like(a)
    m1 = a <= 0.5
    out = np.where(m1, (1.0 + np.sqrt(np.clip(1.0 - 3.0 * a * a, 0.0, None))) / (3.0 * h), out)
    m2 = (a > 0.5) & (a <= 1.5)
    out = np.where(m2, (5.0 - 3.0 * a - np.sqrt(np.clip(1.0 - 3.0 * (1.0 - a) ** 2, 0.0, None))) / (6.0 * h), out)
    return out


class ImmersedBoundary:
    """ib_immersed.py:22-120 (explicit forcing)."""

    def __init__(self, mesh, hh=None):
        idx = np.asarray(mesh.index_cell)
        self.n = np.asarray(idx.shape)
        self.pos_ref = np.asarray(mesh.pos_ref, float)
        self.dx = float(np.asarray(mesh.dx_uniform)[0])
        self.hh = float(self.dx if hh is None else hh)
        self.ncell = mesh.cells.shape[0]
        self.index_cell = idx
        o = np.arange(-2, 3, dtype=np.int64)
        self._offsets = tuple(a.ravel() for a in np.meshgrid(o, o, o, indexing="ij"))
        self._stencil = None
        self._stencil_markers = None
        self._imp = None

    def stencil(self, m):
        """ib_immersed.py:46-61 value-keyed cache + 63-81 build."""
        m = np.asarray(m, float)
        if self._stencil is not None and self._stencil_markers.shape == m.shape \
                and np.array_equal(m, self._stencil_markers):
            return self._stencil
        self._stencil = self.build(m)
        self._stencil_markers = m.copy()
        return self._stencil

    def build(self, m):
        l = np.rint((m - self.pos_ref) / self.dx).astype(np.int64)
        di, dj, dk = self._offsets
        ii = l[:, 0:1] + di[None, :]
        jj = l[:, 1:2] + dj[None, :]
        kk = l[:, 2:3] + dk[None, :]
        nx, ny, nz = (int(v) for v in self.n)
        inb = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
        cell = self.index_cell[np.clip(ii, 0, nx - 1), np.clip(jj, 0, ny - 1), np.clip(kk, 0, nz - 1)]
        valid = inb & (cell >= 0)
        gx = self.pos_ref[0] + ii * self.dx - m[:, 0:1]
        gy = self.pos_ref[1] + jj * self.dx - m[:, 1:2]
        gz = self.pos_ref[2] + kk * self.dx - m[:, 2:3]
        w = delta_h3(gx, self.hh) * delta_h3(gy, self.hh) * delta_h3(gz, self.hh)
        return np.where(valid, cell, 0), np.where(valid, w, 0.0)

    def interp_stencil(self, vc, st):
        cell, w = st
        w = w * self.hh ** 3
        out = np.zeros((cell.shape[0], vc.shape[1]))
        for c in range(vc.shape[1]):
            out[:, c] = (vc[cell, c] * w).sum(axis=1)
        return out

    def spread_stencil(self, flag, st, ds):
        cell, w = st
        w = w * (np.asarray(ds, float) ** 3)[:, None]
        out = np.zeros((self.ncell, flag.shape[1]))
        for c in range(flag.shape[1]):
            out[:, c] = np.bincount(cell.ravel(), (flag[:, c:c + 1] * w).ravel(), minlength=self.ncell)
        return out

    def interp(self, vc, markers):
        return self.interp_stencil(vc, self.stencil(markers))

    def spread(self, flag, markers, ds):
        return self.spread_stencil(flag, self.stencil(markers), ds)

    def direct_force(self, vc, markers, vel_desired, dt, ds):
        st = self.stencil(markers)
        f_lag = (vel_desired - self.interp_stencil(vc, st)) / dt
        return self.spread_stencil(f_lag, st, ds), f_lag

    def prepare_implicit(self, markers, ds, reg=1e-8):
        """ib_immersed.py:122-145: the marker coupling matrix M = D S (D: interpolation weights
        w hh^3, S: spreading weights w ds^3), jittered by reg * trace(M)/n on the diagonal, and its
        dense inverse."""
        import scipy.sparse as sp
        cell, w = self.stencil(markers)
        ds = np.asarray(ds, float)
        w_int = w * self.hh ** 3
        w_spr = w * (ds ** 3)[:, None]
        n = cell.shape[0]
        rows = np.repeat(np.arange(n), cell.shape[1])
        D = sp.csr_matrix((w_int.ravel(), (rows, cell.ravel())), shape=(n, self.ncell))
        S = sp.csr_matrix((w_spr.ravel(), (cell.ravel(), rows)), shape=(self.ncell, n))
        M = np.asarray((D @ S).todense())
        M += reg * np.trace(M) / n * np.eye(n)
        self._imp = dict(cell=cell, w_int=w_int, w_spr=w_spr, M=M, Minv=np.linalg.inv(M), n=n, ds=ds.copy())
        return self

    def invalidate_implicit(self):
        """ib_immersed.py:147-150."""
        self._imp = None

    def implicit_force(self, vc, markers, vel_desired, dt, ds):
        """ib_immersed.py:152-174: f = M^-1 (u_desired - interp(vc))/dt, F_E = spread(f)."""
        self.stencil(markers)
        ds = np.asarray(ds, float)
        if self._imp is None or self._imp["n"] != markers.shape[0] or not np.array_equal(ds, self._imp["ds"]):
            self.prepare_implicit(markers, ds)
        c = self._imp
        cell = c["cell"]
        vel_ib = np.empty((cell.shape[0], vc.shape[1]))
        for k in range(vc.shape[1]):
            vel_ib[:, k] = (vc[cell, k] * c["w_int"]).sum(axis=1)
        f = c["Minv"] @ ((np.asarray(vel_desired, float) - vel_ib) / dt)
        fE = np.zeros((self.ncell, vc.shape[1]))
        for k in range(vc.shape[1]):
            fE[:, k] = np.bincount(cell.ravel(), (f[:, k:k + 1] * c["w_spr"]).ravel(), minlength=self.ncell)
        return fE, f
