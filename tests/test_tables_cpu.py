"""Host-side checks of the product's gather tables (no GPU).

A NumPy emulation of each kernel's gather (one output element at a time,
reading the tables exactly as the HIP kernels do) must reproduce the oracle
operators on a 3-level nested mesh with 2:1 interfaces.
"""
import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.mesh as PM
from pyib_hip.geometry import Geometry, diag_preconditioner
from pyib_hip.tables import build_tables
from oracle.pyib_ref import core as OCore
from oracle.pyib_ref import driver as OD
from tests.helpers import rel


def sgather(x, idx):
    return np.where(idx >= 0, x[np.where(idx >= 0, idx, 0)], -x[np.where(idx < 0, -idx - 1, 0)])


def seg_sum(vals, ptr):
    """Sequential per-row sums (the kernel's loop order)."""
    out = np.zeros((ptr.shape[0] - 1,) + vals.shape[1:])
    for r in range(ptr.shape[0] - 1):
        acc = np.zeros(vals.shape[1:])
        for j in range(ptr[r], ptr[r + 1]):
            acc = acc + vals[j]
        out[r] = acc
    return out


@pytest.fixture(scope="module")
def setup():
    m = PM.make_nested_mesh(0.125, PM.nested_box_boxes([1.5, 1.0, 1.0], 3))
    g = Geometry(m)
    return m, g, build_tables(g), OCore.Core(m)


def test_curl_rot_tables(setup):
    m, g, t, oc = setup
    rng = np.random.default_rng(0)
    e, f = rng.standard_normal(g.nedge), rng.standard_normal(g.nface)
    assert np.array_equal(seg_sum(sgather(e, t.curl_idx), t.curl_ptr), oc.curl(e))
    assert np.array_equal(seg_sum(sgather(f, t.rot_idx), t.rot_ptr), oc.rot(f))


def test_cell_face_tables(setup):
    m, g, t, oc = setup
    rng = np.random.default_rng(1)
    U = rng.standard_normal(g.nface)
    ncell = g.ncell
    first = np.zeros(t.cf_face.shape[0], bool)
    for c in range(ncell):
        first[t.cf_ptr[c]:t.cf_mid[c]] = True
    f = t.cf_face
    a = seg_sum(np.where(first[:, None], U[f, None] * g.rf1[f], 0.0), t.cf_ptr)
    b = seg_sum(np.where(~first[:, None], U[f, None] * g.rf2[f], 0.0), t.cf_ptr)
    vc = (a - b) * g.vol_ci[:, None]
    assert rel(vc, oc.interp_f2c(U)) < 1e-15
    # cdiv with the same split
    a = seg_sum(np.where(first, U[f], 0.0), t.cf_ptr)
    b = seg_sum(np.where(~first, U[f], 0.0), t.cf_ptr)
    assert rel(a - b, oc.cdiv(U)) < 1e-15
    # every first-owner entry's other is the second owner or a ghost index
    cell_of = np.repeat(np.arange(ncell), np.diff(t.cf_ptr))
    assert np.all(g.F2C[f[first], 0] == cell_of[first])
    assert np.all(g.F2C[f[~first], 1] == cell_of[~first])
    ghost = t.cf_other[first] >= ncell
    assert np.array_equal(t.bf[t.cf_other[first][ghost] - ncell], f[first][ghost])


def test_flow_mat_emulated(setup):
    m, g, t, oc = setup
    visc, rou, dt = 0.02, 1.0, 0.01
    rng = np.random.default_rng(2)
    p = rng.standard_normal(g.nedge)
    U = seg_sum(sgather(p, t.curl_idx), t.curl_ptr)
    ncell = g.ncell
    first = np.zeros(t.cf_face.shape[0], bool)
    for c in range(ncell):
        first[t.cf_ptr[c]:t.cf_mid[c]] = True
    f = t.cf_face
    vc = (seg_sum(np.where(first[:, None], U[f, None] * g.rf1[f], 0), t.cf_ptr)
          - seg_sum(np.where(~first[:, None], U[f, None] * g.rf2[f], 0), t.cf_ptr)) * g.vol_ci[:, None]
    own = np.repeat(np.arange(ncell), np.diff(t.cf_ptr))
    o = t.cf_other
    xo = np.where((o < ncell)[:, None], vc[np.minimum(o, ncell - 1)], 0.0)
    k = (g.AoL[f] * visc)[:, None]
    a = seg_sum(np.where(first[:, None], k * (xo - vc[own]), 0), t.cf_ptr)
    b = seg_sum(np.where(~first[:, None], k * (vc[own] - xo), 0), t.cf_ptr)
    v2 = rou * vc - (0.5 * dt * g.vol_ci)[:, None] * (a - b)
    F = (v2[t.f2c[:, 0]] * g.rf1).sum(1)
    i2 = t.f2c[:, 1] >= 0
    F[i2] -= (v2[t.f2c[i2, 1]] * g.rf2[i2]).sum(1)
    w = seg_sum(sgather(F, t.rot_idx), t.rot_ptr)
    ref = oc.flow_mat(p.copy(), alpd=0.5, visc=visc, rou=rou, dt=dt)
    assert rel(w, ref) < 1e-13


def test_preconditioner_matches_oracle(setup):
    m, g, t, oc = setup
    mask = np.isin(np.asarray(m.bc_edge), [3, 4, 5, 6])
    assert np.array_equal(diag_preconditioner(g, mask), OD.diag_preconditioner(oc, mask))
