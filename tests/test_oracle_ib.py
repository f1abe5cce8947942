"""Pins of the oracle's immersed-boundary transfer and kinematics.

Partition of unity, first-moment (linear) reproduction, force conservation
(the pattern of the reference test tests/test_ib_postprocess.py:118-127),
the duality <spread f, v> h^3 = <f ds^3, interp v>, the closed form of the
3-point kernel and analytic-velocity = central-difference for the oscillation.
"""
import numpy as np

from oracle.pyib_ref import immersed as OI
from oracle.pyib_ref import kinematics as OK
from oracle.pyib_ref import mesh as OM
from oracle.pyib_ref import core as OCore


def setup():
    m = OM.make_nested_mesh(0.125, OM.uniform_box([[-1, 1]] * 3))
    return m, OCore.Core(m), OI.ImmersedBoundary(m)


def test_delta_closed_form():
    h = 0.1
    r = np.array([0.0, 0.5, 1.0, 1.5, 2.0]) * h
    expect = np.array([2.0 / 3.0, (1 + np.sqrt(0.25)) / 3, 1.0 / 6.0, 0.0, 0.0]) / h
    assert np.allclose(OI.delta_h3(r, h), expect, rtol=1e-15)
    x = np.linspace(-2, 2, 4001) * h
    assert abs(np.trapezoid(OI.delta_h3(x, h), x) - 1.0) < 1e-6


def test_partition_linear_and_conservation():
    m, c, ib = setup()
    rng = np.random.default_rng(0)
    mk = rng.uniform(-0.5, 0.5, size=(200, 3))
    one = np.ones((c.ncell, 1))
    assert np.abs(ib.interp(one, mk)[:, 0] - 1.0).max() < 1e-13
    for k in range(3):
        assert np.abs(ib.interp(c.cpos[:, k:k + 1], mk)[:, 0] - mk[:, k]).max() < 1e-13
    ds = rng.uniform(0.05, 0.15, 200)
    f = rng.standard_normal((200, 3))
    FE = ib.spread(f, mk, ds)
    assert np.abs(FE.sum(axis=0) * ib.hh ** 3 - (f * ds[:, None] ** 3).sum(axis=0)).max() < 1e-13
    v = rng.standard_normal((c.ncell, 3))
    lhs = (FE * v).sum() * ib.hh ** 3
    rhs = (f * ds[:, None] ** 3 * ib.interp(v, mk)).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_oscillation_velocity_is_derivative():
    rng = np.random.default_rng(0)
    mo = OK.Oscillation(rng.standard_normal((50, 3)), axis=(0, 1, 0), A=0.0625, freq=0.8)
    eps = 1e-6
    for t in (0.0, 0.13, 0.41, 0.77, 1.9):
        fd = (mo(t + eps)[0] - mo(t - eps)[0]) / (2 * eps)
        assert np.abs(fd - mo(t)[1]).max() < 1e-5
