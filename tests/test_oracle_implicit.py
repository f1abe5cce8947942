"""Pins of the oracle's implicit IB forcing (ib_immersed.py:122-174) and rotating-sphere velocity
(ib_case.py:566-584).

* Defining property of the implicit force: the corrected field's interpolated velocity equals
  the desired one, interp(vc + dt F_E) = u_d, up to the 1e-8 diagonal jitter (the explicit force
  leaves the delta-overlap error, which this checks is large).
* The coupling matrix equals a brute-force product of the interpolation and spreading weights
  built cell by cell, and is symmetric for equal marker widths.
* A planted mistake (spreading with the interpolation weights hh^3 instead of ds^3) fails the
  slip check.
* Rotation: omega x (x - c) is tangent, perpendicular to the axis, of size spin U d_axis / R.
"""
import numpy as np
import pytest

from oracle.pyib_ref import case as OC
from oracle.pyib_ref import immersed as OI
from oracle.pyib_ref import mesh as OM


@pytest.fixture(scope="module")
def setup():
    mesh = OM.make_nested_mesh(0.0625, OM.uniform_box([[-0.5, 0.5], [-0.5, 0.5], [-0.5, 0.5]]))
    ib = OI.ImmersedBoundary(mesh)
    n = max(12, int(round(4 * np.pi * 0.25 ** 2 / 0.0625 ** 2)))
    mk = OC.fibonacci_sphere(n, 0.25, [0.01, -0.02, 0.015])
    ds = np.full(n, (4 * np.pi * 0.25 ** 2 * 0.0625 / n) ** (1 / 3))
    rng = np.random.default_rng(2)
    vc = rng.standard_normal((ib.ncell, 3))
    ud = rng.standard_normal((n, 3))
    return ib, mk, ds, vc, ud


def test_implicit_force_removes_the_slip(setup):
    ib, mk, ds, vc, ud = setup
    dt = 0.01
    fE, f = ib.implicit_force(vc, mk, ud, dt, ds)
    slip0 = np.linalg.norm(ib.interp(vc, mk) - ud)
    after = np.linalg.norm(ib.interp(vc + dt * fE, mk) - ud)
    fx, _ = ib.direct_force(vc, mk, ud, dt, ds)
    explicit = np.linalg.norm(ib.interp(vc + dt * fx, mk) - ud)
    assert after <= 1e-6 * slip0
    assert explicit >= 1e-2 * slip0


def test_coupling_matrix_is_the_weight_product(setup):
    ib, mk, ds, _, _ = setup
    ib.prepare_implicit(mk, ds, reg=0.0)
    M = ib._imp["M"]
    cell, w = ib.stencil(mk)
    n = mk.shape[0]
    D = np.zeros((n, ib.ncell))
    S = np.zeros((ib.ncell, n))
    for i in range(n):
        for c, wi in zip(cell[i], w[i]):
            D[i, c] += wi * ib.hh ** 3
            S[c, i] += wi * ds[i] ** 3
    np.testing.assert_allclose(M, D @ S, rtol=1e-12, atol=1e-15 * np.abs(M).max())
    np.testing.assert_allclose(M, M.T, rtol=1e-12, atol=1e-15 * np.abs(M).max())


def test_planted_weight_error_is_caught(setup):
    ib, mk, ds, vc, ud = setup
    dt = 0.01
    fE, f = ib.implicit_force(vc, mk, ud, dt, ds * 1.3)      # forces spread with the wrong widths
    cell, w = ib.stencil(mk)
    wrong = np.zeros_like(vc)
    for k in range(3):
        wrong[:, k] = np.bincount(cell.ravel(), (f[:, k:k + 1] * w * (ds ** 3)[:, None]).ravel(), minlength=ib.ncell)
    assert np.linalg.norm(ib.interp(vc + dt * wrong, mk) - ud) > 1e-3 * np.linalg.norm(ib.interp(vc, mk) - ud)


def test_rotation_velocity_closed_form():
    c = np.array([0.1, -0.2, 0.05])
    mk = OC.fibonacci_sphere(50, 0.25, c)
    axis = np.array([0.0, 2.0, 0.0])
    v = OC.rotating_sphere_velocity(mk, c, axis, 0.5, 0.25, 1.0)
    r = mk - c
    np.testing.assert_allclose((v * r).sum(1), 0.0, atol=1e-15)
    np.testing.assert_allclose(v[:, 1], 0.0, atol=1e-15)
    d_axis = np.hypot(r[:, 0], r[:, 2])
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 0.5 * 1.0 / 0.25 * d_axis, rtol=1e-13)
    with pytest.raises(ValueError):
        OC.rotating_sphere_velocity(mk, c, [0, 0, 0], 0.5, 0.25, 1.0)
