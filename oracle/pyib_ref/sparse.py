"""Oracle restatement of the assembled elliptic operator (SciPy CSR).

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_sparse.py:23-96``:
``A = R @ I @ Vblock @ P @ C`` with ``R = C^T``, canonicalised with
``sum_duplicates`` + ``sort_indices``.  This is the reference's production CG
matvec (``solver.sparse: true``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def curl_matrix(core):
    """ib_sparse.py:23-26."""
    return sp.coo_matrix((core.fe_sign_link, (core.lf_face, core.lf_edge)),
                         shape=(core.nface, core.nedge)).tocsr()


def interp_f2c_matrix(core):
    """ib_sparse.py:29-38 (component-major rows d*ncell + c)."""
    nc, m1, F2C = core.ncell, core.m1, core.F2C
    f = np.arange(core.nface)
    rows, cols, vals = [], [], []
    for d in range(3):
        rows += [d * nc + F2C[:, 0], d * nc + F2C[m1, 1]]
        cols += [f, f[m1]]
        vals += [core.vol_ci[F2C[:, 0]] * core.rf1[:, d], -core.vol_ci[F2C[m1, 1]] * core.rf2[m1, d]]
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(3 * nc, core.nface)).tocsr()


def integr_c2f_matrix(core):
    """ib_sparse.py:41-50."""
    nc, m1, F2C = core.ncell, core.m1, core.F2C
    f = np.arange(core.nface)
    rows, cols, vals = [], [], []
    for d in range(3):
        rows += [f, f[m1]]
        cols += [d * nc + F2C[:, 0], d * nc + F2C[m1, 1]]
        vals += [core.rf1[:, d], -core.rf2[m1, d]]
    return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(core.nface, 3 * nc)).tocsr()


def cdiv_matrix(core):
    """ib_sparse.py:53-58."""
    m1, F2C = core.m1, core.F2C
    rows = np.concatenate([F2C[:, 0], F2C[m1, 1]])
    cols = np.concatenate([np.arange(core.nface), np.arange(core.nface)[m1]])
    vals = np.concatenate([np.ones(core.nface), -np.ones(int(m1.sum()))])
    return sp.coo_matrix((vals, (rows, cols)), shape=(core.ncell, core.nface)).tocsr()


def cgrad_ghost0_matrix(core):
    """ib_sparse.py:61-66."""
    m1, F2C = core.m1, core.F2C
    rows = np.concatenate([np.arange(core.nface), np.arange(core.nface)[m1]])
    cols = np.concatenate([F2C[:, 0], F2C[m1, 1]])
    vals = np.concatenate([-np.ones(core.nface), np.ones(int(m1.sum()))])
    return sp.coo_matrix((vals, (rows, cols)), shape=(core.nface, core.ncell)).tocsr()


def viscous_block(core, visc, rou, dt):
    """ib_sparse.py:79-83: V = rou*I - 0.5*dt*diag(vol_ci) CDIV diag(AoL*visc) CGRAD0."""
    V = rou * sp.eye(core.ncell) - 0.5 * dt * (
        sp.diags(core.vol_ci) @ cdiv_matrix(core) @ sp.diags(core.AoL * visc) @ cgrad_ghost0_matrix(core))
    return V.tocsr()


def build_A(core, alpd=0.0, visc=0.0, rou=1.0, dt=0.0):
    """ib_sparse.py:69-96 (host CSR, float64)."""
    C = curl_matrix(core)
    P = interp_f2c_matrix(core)
    I = integr_c2f_matrix(core)
    R = C.T.tocsr()
    if alpd != 0.0:
        V = viscous_block(core, visc, rou, dt)
        A = (R @ I @ sp.block_diag([V, V, V]).tocsr() @ P @ C).tocsr()
    else:
        A = (R @ I @ P @ C).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A
