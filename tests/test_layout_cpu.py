"""Structured layout of the elliptic operator, checked on the CPU (no GPU).

A NumPy emulation of the structured stencils on the regular rows plus the
composite rows of the irregular ones must reproduce the oracle's flow_mat
(ib_core_np.py:213-220) through the signed renumbering, on 2:1 nested meshes
with dyadic and non-dyadic spacing.
"""
import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.mesh as PM
from pyib_hip.geometry import Geometry
from pyib_hip.layout import build_layout, row_volumes, to_internal_edges, to_reference_edges
from pyib_hip.tables import build_tables
from oracle.pyib_ref import core as OCore
from tests.helpers import rel


def blocks(lay, lv, v):
    shp = [(lv.nx, lv.ny + 1, lv.nz + 1), (lv.nx + 1, lv.ny, lv.nz + 1), (lv.nx + 1, lv.ny + 1, lv.nz)]
    return [v[lv.e_off[d]:lv.e_off[d] + int(np.prod(shp[d]))].reshape(shp[d]) for d in range(3)]


def emulate(lay, p, rou, visc, hdt, cells=False):
    """Structured stencils on regular rows + composite rows elsewhere (the kernels' arithmetic);
    with ``cells`` also the intermediate cell vectors x = interp curl p and V x."""
    vc = np.zeros((3, lay.ncell))
    for lv in lay.levels:
        px, py, pz = blocks(lay, lv, p)
        Ux = py[:, :, :-1] - py[:, :, 1:] + pz[:, 1:, :] - pz[:, :-1, :]
        Uy = px[:, :, 1:] - px[:, :, :-1] - pz[1:, :, :] + pz[:-1, :, :]
        Uz = py[1:, :, :] - py[:-1, :, :] - px[:, 1:, :] + px[:, :-1, :]
        c = 1.0 / (2.0 * lv.h * lv.h)
        v = np.stack([(Ux[:-1] + Ux[1:]) * c, (Uy[:, :-1] + Uy[:, 1:]) * c, (Uz[:, :, :-1] + Uz[:, :, 1:]) * c])
        sl = slice(lv.cell_off, lv.cell_off + lv.ncell)
        reg = lay.reg_g1[sl].astype(bool)
        for d in range(3):
            vc[d, sl][reg] = v[d].ravel()[reg]
    g1 = lay.g1
    for r in range(g1.rows.shape[0]):
        a, b = g1.ptr[r], g1.ptr[r + 1]
        vc[:, g1.rows[r]] = (g1.coef[a:b] * p[g1.col[a:b]][:, None]).sum(axis=0)
    v2 = np.zeros_like(vc)
    for lv in lay.levels:
        sl = slice(lv.cell_off, lv.cell_off + lv.ncell)
        reg = lay.reg_cell[sl].astype(bool)
        k = hdt * visc / (lv.h * lv.h)
        for d in range(3):
            x = np.zeros((lv.nx + 2, lv.ny + 2, lv.nz + 2))
            x[1:-1, 1:-1, 1:-1] = vc[d, sl].reshape(lv.nx, lv.ny, lv.nz)
            lap = (x[2:, 1:-1, 1:-1] + x[:-2, 1:-1, 1:-1] + x[1:-1, 2:, 1:-1] + x[1:-1, :-2, 1:-1]
                   + x[1:-1, 1:-1, 2:] + x[1:-1, 1:-1, :-2] - 6 * x[1:-1, 1:-1, 1:-1])
            y = rou * x[1:-1, 1:-1, 1:-1] - k * lap
            v2[d, sl][reg] = y.ravel()[reg]
    vi = lay.vis
    for r in range(vi.rows.shape[0]):
        a, b = vi.ptr[r], vi.ptr[r + 1]
        c = vi.rows[r]
        v2[:, c] = rou * vc[:, c] - hdt * visc * (vi.coef[a:b][None, :] * vc[:, vi.col[a:b]]).sum(axis=1)
    w = np.zeros(lay.nedge)
    for lv in lay.levels:
        sl = slice(lv.cell_off, lv.cell_off + lv.ncell)
        V = [v2[d, sl].reshape(lv.nx, lv.ny, lv.nz) for d in range(3)]
        hh = 0.5 * lv.h
        shp = [(lv.nx, lv.ny + 1, lv.nz + 1), (lv.nx + 1, lv.ny, lv.nz + 1), (lv.nx + 1, lv.ny + 1, lv.nz)]
        for d in range(3):
            seg = slice(lv.e_off[d], lv.e_off[d] + int(np.prod(shp[d])))
            idx = np.flatnonzero(lay.reg_edge[seg])
            i, j, k = np.unravel_index(idx, shp[d])
            Fx = lambda i, j, k: hh * (V[0][i - 1, j, k] + V[0][i, j, k])
            Fy = lambda i, j, k: hh * (V[1][i, j - 1, k] + V[1][i, j, k])
            Fz = lambda i, j, k: hh * (V[2][i, j, k - 1] + V[2][i, j, k])
            if d == 0:
                val = -Fy(i, j, k) + Fy(i, j, k - 1) + Fz(i, j, k) - Fz(i, j - 1, k)
            elif d == 1:
                val = Fx(i, j, k) - Fx(i, j, k - 1) - Fz(i, j, k) + Fz(i - 1, j, k)
            else:
                val = -Fx(i, j, k) + Fx(i, j - 1, k) + Fy(i, j, k) - Fy(i - 1, j, k)
            w[seg][idx] = val
    g3 = lay.g3
    for r in range(g3.rows.shape[0]):
        a, b = g3.ptr[r], g3.ptr[r + 1]
        w[g3.rows[r]] = (g3.coef[a:b] * v2[:, g3.col[a:b]].T).sum()
    return (w, vc, v2) if cells else w


@pytest.mark.parametrize("h,boxes", [(0.0625, "nested3"), (0.5 / 24, "nested2"), (0.125, "wake")])
def test_structured_emulation_equals_flow_mat(h, boxes):
    bx = {"nested3": PM.nested_box_boxes([1.5, 1.25, 1.25], 3), "nested2": PM.nested_box_boxes([1.0, 0.75, 0.75], 2),
          "wake": PM.wake_boxes()}[boxes]
    m = PM.make_nested_mesh(h, bx)
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    assert lay.stats["regular_cells"] > 0
    oc = OCore.Core(m)
    rng = np.random.default_rng(7)
    p_ref = rng.standard_normal(g.nedge)
    p_int = to_internal_edges(lay, p_ref)
    kw = dict(visc=0.45, rou=1.0, dt=0.004)
    w_int = emulate(lay, p_int, kw["rou"], kw["visc"], 0.5 * kw["dt"])
    w_ref = oc.flow_mat(p_ref.copy(), alpd=0.5, **kw)
    assert rel(to_reference_edges(lay, w_int), w_ref) < 1e-13
    empty = lay.edge_inv < 0
    assert np.all(w_int[empty] == 0.0)


def test_round_trip_numbering():
    m = PM.make_nested_mesh(0.125, PM.concentric_boxes())
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    x = np.random.default_rng(0).standard_normal(g.nedge)
    assert np.array_equal(to_reference_edges(lay, to_internal_edges(lay, x)), x)
    # level-0 cells keep PyIB's numbering (index_cell is the identity there)
    n0 = m.layer_cell_counts[0]
    assert np.array_equal(lay.cell_map[:n0], np.arange(n0))


def test_sliced_ell_keeps_every_row_sum():
    """The sliced-ELL copy of each composite gives the CSR row sums bit for bit."""
    from pyib_hip.layout import sell_apply, to_sell
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], 2))
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    x = np.random.default_rng(3).standard_normal(max(lay.ncell, lay.nedge))
    for comp in (lay.g1, lay.vis, lay.g3):
        sc = to_sell(comp)
        assert sc.rows.shape[0] % 64 == 0 and np.all(sc.col[sc.coef.reshape(sc.col.shape[0], -1).any(axis=1) == 0]
                                                     >= 0)
        got = sell_apply(sc, lambda c: x[c][:, None] if comp.coef.ndim == 2 else x[c])
        assert sorted(got) == sorted(int(r) for r in comp.rows)
        for k, r in enumerate(comp.rows):
            a, b = comp.ptr[k], comp.ptr[k + 1]
            terms = comp.coef[a:b] * (x[comp.col[a:b]][:, None] if comp.coef.ndim == 2 else x[comp.col[a:b]])
            assert np.array_equal(got[int(r)], np.cumsum(terms, axis=0)[-1])      # sequential, as the kernels


@pytest.mark.parametrize("h,boxes", [(0.0625, "nested3"), (0.125, "wake")])
def test_quadratic_form_from_cell_passes(h, boxes):
    """p.A p as the V passes form it (h^3 x.(V x) on regular cells, row_volumes-weighted on
    the composite rows) equals p.w, and the oracle's p.(A p), when p is zero off the real
    edges; a wrong weight (vol_ci instead of its inverse) or a cell left out breaks it."""
    bx = {"nested3": PM.nested_box_boxes([1.5, 1.25, 1.25], 3), "wake": PM.wake_boxes()}[boxes]
    m = PM.make_nested_mesh(h, bx)
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    rng = np.random.default_rng(11)
    p_ref = rng.standard_normal(g.nedge)
    p = to_internal_edges(lay, p_ref)
    rou, visc, dt = 1.0, 0.45, 0.004
    w, vc, y = emulate(lay, p, rou, visc, 0.5 * dt, cells=True)
    wt = np.zeros(lay.ncell)
    for lv in lay.levels:
        sl = slice(lv.cell_off, lv.cell_off + lv.ncell)
        wt[sl][lay.reg_cell[sl].astype(bool)] = lv.h ** 3
    wt[lay.vis.rows] = row_volumes(lay, g, lay.vis.rows)
    quad = float((wt * (vc * y).sum(axis=0)).sum())
    pw = float(p @ w)
    ref = float(p_ref @ OCore.Core(m).flow_mat(p_ref.copy(), alpd=0.5, visc=visc, rou=rou, dt=dt))
    assert abs(quad - pw) <= 1e-13 * abs(pw)
    assert abs(quad - ref) <= 1e-13 * abs(ref)
    assert np.all(row_volumes(lay, g, np.array([-1])) == 0.0)
