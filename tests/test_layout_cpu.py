"""Structured layout of the elliptic operator, checked on the CPU (no GPU).

A NumPy emulation of the structured stencils on the regular rows plus the
composite rows of the irregular ones must reproduce the oracle's flow_mat
(ib_core_np.py:168-175) through the signed renumbering, on 2:1 nested meshes
with dyadic and non-dyadic spacing.
"""
import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.mesh as PM
from pyib_hip.geometry import Geometry
from pyib_hip.layout import (build_layout, composite_rows, g1v_flags, row_volumes, to_internal_edges,
                             to_reference_edges)
from pyib_hip.tables import build_tables
from oracle.pyib_ref import core as OCore
from tests.helpers import rel


def blocks(lay, lv, v):
    shp = [(lv.nx, lv.ny + 1, lv.nz + 1), (lv.nx + 1, lv.ny, lv.nz + 1), (lv.nx + 1, lv.ny + 1, lv.nz)]
    return [v[lv.e_off[d]:lv.e_off[d] + int(np.prod(shp[d]))].reshape(shp[d]) for d in range(3)]


def box_faces(lv):
    """Number of each cell's faces on the level box's boundary."""
    i, j, k = np.meshgrid(np.arange(lv.nx), np.arange(lv.ny), np.arange(lv.nz), indexing="ij")
    return ((i == 0).astype(int) + (i == lv.nx - 1) + (j == 0) + (j == lv.ny - 1) + (k == 0) + (k == lv.nz - 1))


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
        nbox = box_faces(lv) * ((lay.reg_cell[sl] & 2) != 0).reshape(lv.nx, lv.ny, lv.nz)
        k = hdt * visc / (lv.h * lv.h)
        for d in range(3):
            x = np.zeros((lv.nx + 2, lv.ny + 2, lv.nz + 2))
            x[1:-1, 1:-1, 1:-1] = vc[d, sl].reshape(lv.nx, lv.ny, lv.nz)
            lap = (x[2:, 1:-1, 1:-1] + x[:-2, 1:-1, 1:-1] + x[1:-1, 2:, 1:-1] + x[1:-1, :-2, 1:-1]
                   + x[1:-1, 1:-1, 2:] + x[1:-1, 1:-1, :-2] - 6 * x[1:-1, 1:-1, 1:-1])
            # a face that leaves the box (zero ghost) has AoL = 2h: twice the coefficient
            lap = lap - nbox * x[1:-1, 1:-1, 1:-1]
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


def test_g1v_flags_cover_the_composite_v_reads():
    """The fused G1+V pass stores x only where bit 1 is set; every cell a composite V row
    reads (its own row and its columns) must then be stored or come from the composite G1
    rows (not a regular G1 row)."""
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], 3))
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    f = g1v_flags(lay)
    assert np.array_equal(f & 1, lay.reg_g1)
    assert np.array_equal((f >> 2) & 1, lay.reg_cell & 1)
    assert np.array_equal((f >> 4) & 1, (lay.reg_cell >> 1) & 1)
    assert not np.any((f & 2).astype(bool) & ~(f & 1).astype(bool))
    read = np.unique(np.concatenate([lay.vis.rows, lay.vis.col]))
    ok = ((f[read] & 2) != 0) | (lay.reg_g1[read] == 0)
    assert ok.all()
    assert 0 < int((f & 2).astype(bool).sum()) < int(lay.reg_g1.sum())


# ------------------------------------------------------------- fused G1+V tile loop
TJ, TK = 8, 32                                   # tile of k_sg1v_t
RJ, RK = TJ + 2, TK + 2                          # cells with the one-cell halo
XJ, XK, YJ, YK, ZJ, ZK = RJ + 1, RK + 1, RJ, RK + 1, RJ + 1, RK


def _tile(lay, lv, p, flags, xmem, y, j0, k0, i0, i1, rou, visc, hdt):
    nx, ny, nz, co = lv.nx, lv.ny, lv.nz, lv.cell_off
    e0, e1, e2 = lv.e_off
    ny1, nz1 = ny + 1, nz + 1
    cf, kA, hv = 1.0 / (2 * lv.h * lv.h), lv.h * visc, hdt / lv.h ** 3
    jb, kb = j0 - 1, k0 - 1

    def edges(d, ip):
        """Staged edge plane (0 outside the level's edge blocks), as the kernel's fetch."""
        J, K = {0: (XJ, XK), 1: (YJ, YK), 2: (ZJ, ZK)}[d]
        q = np.arange(J * K)
        j, k = jb + q // K, kb + q % K
        nj, nk, ni = {0: (ny1, nz1, nx), 1: (ny, nz1, nx + 1), 2: (ny1, nz, nx + 1)}[d]
        ok = (ip >= 0) & (ip < ni) & (j >= 0) & (j < nj) & (k >= 0) & (k < nk)
        idx = (e0, e1, e2)[d] + (ip * nj + j) * nk + k
        return np.where(ok, p[np.where(ok, idx, 0)], 0.0)

    def cell_flags(ip):
        q = np.arange(RJ * RK)
        j, k = jb + q // RK, kb + q % RK
        ok = (ip >= 0) & (ip < nx) & (j >= 0) & (j < ny) & (k >= 0) & (k < nz)
        return np.where(ok, 0x80 | flags[np.where(ok, co + (ip * ny + j) * nz + k, 0)], 0)

    slot = lambda i: (i + 2) & 1              # noqa: E731
    sy, sz, sc = [None, None], [None, None], [None, None]
    sy[slot(i0 - 1)], sz[slot(i0 - 1)] = edges(1, i0 - 1), edges(2, i0 - 1)
    tj, tk = (v.ravel() for v in np.meshgrid(np.arange(TJ), np.arange(TK), indexing="ij"))
    qo = (tj + 1) * RK + tk + 1
    xm = np.zeros((3, TJ * TK))
    q = np.arange(RJ * RK)
    jj, kk = q // RK, q % RK
    for ip in range(i0 - 1, i1 + 1):
        a, b = slot(ip), slot(ip + 1)
        sx, sy[b], sz[b] = edges(0, ip), edges(1, ip + 1), edges(2, ip + 1)
        fc = cell_flags(ip)
        PX = lambda dj, dk: sx[(jj + dj) * XK + kk + dk]      # noqa: E731
        PY = lambda s, dk: sy[s][jj * YK + kk + dk]           # noqa: E731
        PZ = lambda s, dj: sz[s][(jj + dj) * ZK + kk]         # noqa: E731
        ux = (PY(a, 0) - PY(a, 1)) + (PZ(a, 1) - PZ(a, 0)) + ((PY(b, 0) - PY(b, 1)) + (PZ(b, 1) - PZ(b, 0)))
        uy = (PX(0, 1) - PX(0, 0)) - (PZ(b, 0) - PZ(a, 0)) + ((PX(1, 1) - PX(1, 0)) - (PZ(b, 1) - PZ(a, 1)))
        uz = (PY(b, 0) - PY(a, 0)) - (PX(1, 0) - PX(0, 0)) + ((PY(b, 1) - PY(a, 1)) - (PX(1, 1) - PX(0, 1)))
        comp = np.stack([ux * cf, uy * cf, uz * cf])
        c = co + (ip * ny + (jb + jj)) * nz + (kb + kk)
        reg, inb = (fc & 1) != 0, (fc & 0x80) != 0
        mem = np.stack([xmem[3 * np.where(inb, c, 0) + d] for d in range(3)])
        own = (jj >= 1) & (jj <= TJ) & (kk >= 1) & (kk <= TK) & (ip >= i0) & (ip < i1)
        wr = reg & ((fc & 2) != 0) & own
        for d in range(3):
            xmem[3 * c[wr] + d] = comp[d][wr]
        sc[a] = np.where(reg, comp, np.where(inb, mem, 0.0))
        i = ip - 1
        if i >= i0:
            j, k = j0 + tj, k0 + tk
            ok = (j < ny) & (k < nz)
            cc = co + (i * ny + j) * nz + k
            fl = flags[np.where(ok, cc, 0)]
            rv = ok & ((fl & 4) != 0)
            nbox = np.where((fl & 16) != 0, ((i == 0) + (i == nx - 1) + (j == 0) + (j == ny - 1) + (k == 0)
                                              + (k == nz - 1)).astype(float), 0.0)
            s0 = slot(i)
            for d in range(3):
                xc = sc[s0][d][qo]
                # the face coefficient kA factored out of the six faces, as the kernel does
                aa = ((sc[a][d][qo] - xc) + (sc[s0][d][qo + RK] - xc)) + (sc[s0][d][qo + 1] - xc)
                aa = aa - nbox * xc
                bb = ((xc - xm[d]) + (xc - sc[s0][d][qo - RK])) + (xc - sc[s0][d][qo - 1])
                y[3 * cc[rv] + d] = (rou * xc - (hv * kA) * (aa - bb))[rv]
        if ip >= i0:
            xm = np.stack([sc[slot(ip - 1)][d][qo] for d in range(3)])


def emulate_fused_g1v(lay, p, rou, visc, hdt, ci=16):
    """The fused G1+V tile loop of k_sg1v_t restated in NumPy: same tiles, halo, plane ring,
    register-kept plane i-1 and store rule.  Returns (y, x as stored in HBM)."""
    flags = g1v_flags(lay)
    xmem = np.zeros(3 * lay.ncell)
    g1 = lay.g1                                  # composite G1 rows run first
    for r in range(g1.rows.shape[0]):
        a, b = g1.ptr[r], g1.ptr[r + 1]
        xmem[3 * g1.rows[r]:3 * g1.rows[r] + 3] = (g1.coef[a:b] * p[g1.col[a:b]][:, None]).sum(axis=0)
    y = np.zeros(3 * lay.ncell)
    for lv in lay.levels:
        for j0 in range(0, lv.ny, TJ):
            for k0 in range(0, lv.nz, TK):
                for i0 in range(0, lv.nx, ci):
                    _tile(lay, lv, p, flags, xmem, y, j0, k0, i0, min(i0 + ci, lv.nx), rou, visc, hdt)
    return y, xmem


@pytest.mark.parametrize("ci", [4, 16])
def test_fused_g1v_tile_loop_equals_two_passes(ci):
    """Index math of the fused G1+V pass (halo, plane ring, i-chunks) against the two-pass
    emulation: y on every regular V cell and x on every stored cell."""
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], 3))
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    rng = np.random.default_rng(5)
    p = rng.standard_normal(lay.nedge)
    p[lay.edge_inv < 0] = 0.0
    _, vc, v2 = emulate(lay, p, 1.0, 0.45, 0.002, cells=True)
    y, xmem = emulate_fused_g1v(lay, p, 1.0, 0.45, 0.002, ci=ci)
    regv = lay.reg_cell.astype(bool)
    yy = y.reshape(-1, 3).T
    assert np.abs(yy[:, regv] - v2[:, regv]).max() <= 1e-14 * np.abs(v2).max()
    keep = (g1v_flags(lay) & 2) != 0
    xx = xmem.reshape(-1, 3).T
    assert np.abs(xx[:, keep] - vc[:, keep]).max() <= 1e-14 * np.abs(vc).max()


def test_composite_rows_keeps_each_selected_row():
    """Dropping rows (the Dirichlet rows of the G3 composite) keeps every other row's entries."""
    m = PM.make_nested_mesh(0.125, PM.concentric_boxes())
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    c = lay.g3
    sel = np.flatnonzero(np.arange(c.rows.shape[0]) % 3 != 1)
    sub = composite_rows(c, sel)
    assert np.array_equal(sub.rows, c.rows[sel])
    for i, r in enumerate(sel[:200]):
        a, b = c.ptr[r], c.ptr[r + 1]
        assert np.array_equal(sub.col[sub.ptr[i]:sub.ptr[i + 1]], c.col[a:b])
        assert np.array_equal(sub.coef[sub.ptr[i]:sub.ptr[i + 1]], c.coef[a:b])
    assert int(sub.ptr[-1]) == int(sum(c.ptr[r + 1] - c.ptr[r] for r in sel))


@pytest.mark.parametrize("levels", [2, 3])
def test_preconditioner_classes(levels):
    """The per-level table of the Jacobi preconditioner: bit 2 only where an edge's own inverse
    diagonal equals the table bitwise; the table is the closed form h/4 (four whole h x h faces,
    1/h each); every regular G3 row carries it; live = neither an empty slot nor Dirichlet."""
    from pyib_hip.geometry import diag_preconditioner
    from pyib_hip.layout import MAX_LEVELS, dir_mask_internal, edge_classes
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], levels))
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    dm = np.asarray(m.bc_edge) > 0
    pint = np.zeros(lay.nedge)
    pint[lay.edge_map] = diag_preconditioner(g, dm)
    dint = dir_mask_internal(lay, dm)
    cls, plev, pg3 = edge_classes(lay, dint, pint)
    assert cls.dtype == np.uint8 and plev.shape == pg3.shape == (MAX_LEVELS,)
    live = (cls & 2) > 0
    assert np.array_equal(live, dint == 0)
    assert not live[lay.edge_inv < 0].any() and not live[lay.edge_map[dm]].any()
    assert np.array_equal(cls & 1, lay.reg_edge)
    lvl = (cls >> 4) & 7
    bounds = [lv.e_off[0] for lv in lay.levels] + [lay.nedge]
    closed = (cls & 4) > 0
    assert not closed[~live].any()
    for L, lv in enumerate(lay.levels):
        assert np.all(lvl[bounds[L]:bounds[L + 1]] == L)
        here = closed & (lvl == L)
        assert np.all(pint[here] == plev[L])
        assert plev[L] == pytest.approx(lv.h / 4.0, rel=1e-14, abs=0)
        assert pg3[L] == plev[L]
        reg = (cls & 1).astype(bool) & (lvl == L)
        assert np.all(closed[reg])
        # the slots left to read their prec are the live edges whose value differs
        rest = live & ~closed & (lvl == L)
        assert np.all(pint[rest] != plev[L])
        assert here.sum() > 0.8 * (live & (lvl == L)).sum()
    # a planted prec value on one regular edge takes bit 2 away from it and the G3 table off
    e = int(np.flatnonzero((cls & 1) & closed)[0])
    bad = pint.copy()
    bad[e] = np.nextafter(bad[e], 1.0)
    cls2, plev2, pg32 = edge_classes(lay, dint, bad)
    L = int((cls2[e] >> 4) & 7)
    assert not (cls2[e] & 4) and pg32[L] == 0.0 and plev2[L] == plev[L]


def test_hole_cells_own_no_regular_row():
    """The tiled kernels skip a tile whose cells all lie in the level's hole: none of those cells
    is a regular G1 or V row or has any flag of the fused pass, and none of the x-, y- and
    z-edges (i, j, k) they own is a regular G3 row or a real edge."""
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], 3))
    g = Geometry(m)
    lay = build_layout(m, g, build_tables(g))
    flags = g1v_flags(lay)
    real_e = np.zeros(lay.nedge, bool)
    real_e[lay.edge_map] = True
    holes = 0
    for lv in lay.levels:
        a, b = lv.hole_lo, lv.hole_hi
        if not b[0] > a[0]:
            continue
        holes += 1
        cells = np.zeros((lv.nx, lv.ny, lv.nz), bool)
        cells[a[0]:b[0], a[1]:b[1], a[2]:b[2]] = True
        sl = slice(lv.cell_off, lv.cell_off + lv.ncell)
        assert not lay.reg_cell[sl][cells.ravel()].any() and not lay.reg_g1[sl][cells.ravel()].any()
        assert not flags[sl][cells.ravel()].any()
        for d, blk in enumerate(blocks(lay, lv, np.arange(lay.nedge))):
            own = blk[:lv.nx, :lv.ny, :lv.nz][cells]
            assert not lay.reg_edge[own].any() and not real_e[own].any()
    assert holes == 2
    # bit 3 (x read from memory) marks exactly the real cells whose G1 row is composite
    real = np.zeros(lay.ncell, bool)
    real[lay.cell_map] = True
    assert np.array_equal((flags & 8) > 0, real & (lay.reg_g1 == 0))
    assert set(np.flatnonzero((flags & 8) > 0)) <= set(np.asarray(lay.g1.rows).tolist())


@pytest.mark.parametrize("h, boxes, kw", [
    (0.0625, "nested3", {}), (0.5 / 24, "nested2", {}), (0.125, "wake", {}), (0.125, "concentric", {}),
    (0.0625, "nested3", {"tol": 0.0}), (0.0625, "nested3", {"margin": 2, "demote_neighbours": True}),
])
def test_native_layout_equals_the_numpy_statement(h, boxes, kw):
    """The C++ builder (pyib_layout_build) produces every array of the NumPy statement
    bit for bit: numbering, orientation, per-pass regularity, and the composite rows with
    their duplicate entries summed in NumPy's add.reduceat order."""
    from pyib_hip.layout import build_layout_numpy
    bx = {"nested3": PM.nested_box_boxes([1.5, 1.25, 1.25], 3), "nested2": PM.nested_box_boxes([1.0, 0.75, 0.75], 2),
          "wake": PM.wake_boxes(), "concentric": PM.concentric_boxes()}[boxes]
    m = PM.make_nested_mesh(h, bx)
    geo = Geometry(m)
    t = build_tables(geo)
    a = build_layout(m, geo, t, **kw)
    b = build_layout_numpy(m, geo, t, **kw)
    for name in ("cell_map", "edge_map", "edge_sigma", "edge_inv", "sigma_int", "reg_cell", "reg_edge", "reg_g1"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.dtype == y.dtype and np.array_equal(x, y), name
    for comp in ("g1", "vis", "g3"):
        x, y = getattr(a, comp), getattr(b, comp)
        for f in ("rows", "ptr", "col", "coef"):
            u, v = getattr(x, f), getattr(y, f)
            assert u.dtype == v.dtype and u.shape == v.shape and np.array_equal(u, v), (comp, f)
    assert a.stats == b.stats
    assert (a.ncell, a.nedge) == (b.ncell, b.nedge)
    for la, lb in zip(a.levels, b.levels):
        assert (la.nx, la.ny, la.nz, la.h, la.lo, la.hole_lo, la.hole_hi, la.cell_off, la.e_off) == \
               (lb.nx, lb.ny, lb.nz, lb.h, lb.lo, lb.hole_lo, lb.hole_hi, lb.cell_off, lb.e_off)


def test_numpy_reduceat_order_is_first_plus_pairwise():
    """The summation order the C++ builder restates: add.reduceat over a run gives the first
    entry plus NumPy's pairwise sum of the rest (checked on runs chosen so that a sequential or
    plain pairwise order gives a different answer)."""
    a = np.array([1.0, 1e16, -1e16])
    assert np.add.reduceat(a, [0])[0] == 1.0 and np.add.reduce(a) == 0.0
    rng = np.random.default_rng(3)
    for n in (2, 7, 8, 9, 16, 17, 127, 129, 300):
        v = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
        rest = v[1:]
        if rest.shape[0] < 8:
            p = 0.0
            for x in rest:
                p += x
        else:
            p = np.add.reduce(rest)          # NumPy's own pairwise sum of a contiguous run
        assert np.add.reduceat(v, [0])[0] == v[0] + p
