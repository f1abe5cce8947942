"""Pins of the oracle's geometry, incidence operators, CSR assembly and CG.

* golden interface geometry and flow_mat written by the reference itself
  (tests/golden/geometry_2level.npz, flow_mat_2level.npz);
* closed forms on uniform regions (Af = h^2, vol_ci = 1/h^3, AoL = h);
* rot is the transpose of curl; A = CSR assembly equals matrix-free flow_mat;
  A is symmetric; A on a 2^3-cell box equals a dense brute-force product of the
  operator chain; CG reaches the dense solution; grad-recon of a linear field.
"""
import numpy as np
import pytest

from oracle.pyib_ref import core as OCore
from oracle.pyib_ref import driver as OD
from oracle.pyib_ref import mesh as OM
from oracle.pyib_ref import sparse as OS
from tests.helpers import golden, rel


def two_level():
    return OM.make_nested_mesh(0.125, [OM.Box.bounds(-0.5, 0.5, -0.5, 0.5, -0.5, 0.5),
                                       OM.Box.bounds(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)])


@pytest.fixture(scope="module")
def core2():
    return OCore.Core(two_level())


def test_geometry_matches_reference_golden(core2):
    with golden("geometry_2level.npz") as g:
        for name in ("fpos", "cpos", "rf", "rf1", "rf2", "nf", "Af", "vol_ci", "AoL", "fe_sign_link"):
            assert np.array_equal(getattr(core2, name), g[name]), name


def test_flow_mat_matches_reference_golden(core2):
    with golden("flow_mat_2level.npz") as g:
        kw = dict(alpd=float(g["alpd"]), visc=float(g["visc"]), rou=float(g["rou"]), dt=float(g["dt"]))
        w = core2.flow_mat(g["p"].copy(), **kw)
        assert np.array_equal(w, g["w_flow_mat"])
        A = OS.build_A(core2, **kw)
        assert rel(A @ g["p"], g["w_csr"]) < 1e-15


def test_uniform_closed_forms():
    h = 0.25
    c = OCore.Core(OM.make_nested_mesh(h, OM.uniform_box([[-1, 1], [-1, 1], [-1, 1]])))
    assert np.allclose(c.Af, h * h, rtol=1e-14)
    assert np.allclose(c.vol_ci, 1.0 / h ** 3, rtol=1e-13)
    assert np.allclose(np.abs(c.AoL[c.m1]), h, rtol=1e-13)
    assert np.allclose(np.abs(c.AoL[~c.m1]), 2 * h, rtol=1e-13)   # boundary: rf = half a cell
    assert np.all(np.abs(c.fe_sign_link) == 1)


def test_interface_volumes_and_centres(core2):
    """After Modify_geom every cell centre is its geometric centre and 1/vol_ci its volume."""
    vol = 1.0 / core2.vol_ci
    n0 = 8 ** 3
    assert np.allclose(vol[:n0], 0.125 ** 3, rtol=1e-12)
    assert np.allclose(vol[n0:], 0.25 ** 3, rtol=1e-12)
    h = np.where(np.arange(core2.ncell) < n0, 0.125, 0.25)
    frac = (core2.cpos + 1.0) / h[:, None] - 0.5
    assert np.allclose(frac, np.rint(frac), atol=1e-9)


def test_rot_is_curl_transpose(core2):
    rng = np.random.default_rng(0)
    e, f = rng.standard_normal(core2.nedge), rng.standard_normal(core2.nface)
    assert abs(f @ core2.curl(e) - e @ core2.rot(f)) < 1e-12 * np.abs(f).sum() * np.abs(e).max()


@pytest.mark.parametrize("alpd", [0.0, 0.5])
def test_csr_equals_flow_mat_and_symmetric(core2, alpd):
    kw = dict(alpd=alpd, visc=0.02, rou=1.0, dt=0.01)
    A = OS.build_A(core2, **kw)
    rng = np.random.default_rng(1)
    p = rng.standard_normal(core2.nedge)
    assert rel(A @ p, core2.flow_mat(p.copy(), **kw)) < 1e-13
    S = (A - A.T).tocoo()
    assert S.data.size == 0 or np.abs(S.data).max() < 1e-12 * np.abs(A.data).max()


def test_brute_force_dense_chain():
    """A on a 2x2x2 box equals the dense product rot . integr . V . interp . curl built column by column."""
    c = OCore.Core(OM.make_nested_mesh(0.5, OM.uniform_box([[-0.5, 0.5]] * 3)))
    kw = dict(alpd=0.5, visc=0.3, rou=1.0, dt=0.1)
    A = OS.build_A(c, **kw).toarray()
    n = c.nedge
    dense = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        # explicit per-cell loops, independent of the vectorised operators
        U = np.zeros(c.nface)
        for k in range(c.lf_face.shape[0]):
            U[c.lf_face[k]] += c.fe_sign_link[k] * e[c.lf_edge[k]]
        vc = np.zeros((c.ncell, 3))
        for f in range(c.nface):
            vc[c.F2C[f, 0]] += U[f] * c.rf1[f]
            if c.F2C[f, 1] >= 0:
                vc[c.F2C[f, 1]] -= U[f] * c.rf2[f]
        vc *= c.vol_ci[:, None]
        lap = np.zeros((c.ncell, 3))
        for f in range(c.nface):
            a, b = c.F2C[f]
            jump = (vc[b] if b >= 0 else 0.0) - vc[a]
            q = c.AoL[f] * kw["visc"] * jump
            lap[a] += q
            if b >= 0:
                lap[b] -= q
        v2 = kw["rou"] * vc - 0.5 * kw["dt"] * c.vol_ci[:, None] * lap
        F = np.array([v2[c.F2C[f, 0]] @ c.rf1[f] - (v2[c.F2C[f, 1]] @ c.rf2[f] if c.F2C[f, 1] >= 0 else 0.0)
                      for f in range(c.nface)])
        w = np.zeros(n)
        for k in range(c.lf_face.shape[0]):
            w[c.lf_edge[k]] += c.fe_sign_link[k] * F[c.lf_face[k]]
        dense[:, j] = w
    assert np.abs(dense - A).max() < 1e-12 * np.abs(A).max()


def test_cg_reaches_dense_solution():
    m = OM.make_nested_mesh(0.25, OM.uniform_box([[-0.5, 0.5]] * 3))
    c = OCore.Core(m)
    drv = OD.Driver(c, m, visc=0.05, cg_max_iterations=5000, cg_check_interval=1)
    drv.set_cavity(U_lid=0.0, spanwise_zerograd=False)
    dt = 0.05
    drv.enable_sparse(dt)
    A = drv.Asp.toarray()
    free = ~drv.dir_mask
    rng = np.random.default_rng(3)
    xt = np.zeros(c.nedge)
    xt[free] = rng.standard_normal(int(free.sum()))
    b = A @ xt                    # manufactured right-hand side in range(A) (A has a gauge null space)
    b[~free] = 0.0
    x, m_it = drv.helm_solve(b.copy(), np.zeros(c.nedge), dt, tol1=1e-14, tol2=1e-13)
    Aff = A[np.ix_(free, free)]
    # A has a gauge null space; compare residuals and the solution projected on range(A)
    assert np.linalg.norm(Aff @ x[free] - b[free]) < 1e-10 * np.linalg.norm(b)
    xs = np.linalg.lstsq(Aff, b[free], rcond=None)[0]
    assert np.linalg.norm(Aff @ (x[free] - xs)) < 1e-10 * np.linalg.norm(b)


def test_grad_recon_linear_exact(core2):
    """ib_core_np.py:186-209 self-test: the gradient of a linear field is e_k."""
    c = core2
    for k in range(3):
        phi = c.cpos[:, k]
        jump = np.empty(c.nface)
        jump[c.m1] = phi[c.F2C[c.m1, 1]] - phi[c.F2C[c.m1, 0]]
        jump[~c.m1] = c.fpos[~c.m1, k] - phi[c.F2C[~c.m1, 0]]
        g = c.grad_recon(jump)
        ek = np.zeros(3)
        ek[k] = 1.0
        assert np.abs(g[c.vol_ci > 0] - ek).max() < 1e-9
