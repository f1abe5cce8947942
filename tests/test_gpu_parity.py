"""HIP path vs the CPU oracle and the reference's golden runs (needs a GPU).

Tolerances (float64 path): incidence operators 1e-14 relative L2, elliptic
operator 1e-13, Helmholtz CG identical iteration count, full steps identical
CG counts and 1e-10 on forces.  Solution fields (CG solution, s, vorticity)
are held to max(1e-12, 3x floor), where floor is the difference between the
reference's own two elliptic paths (CSR matrix of solver.sparse vs the
matrix-free flow_mat) on the same case, computed alongside: a 1000-iteration
CG amplifies round-off, and the reference itself does not reproduce its own
solution closer than that.  Each check prints its error, the floor and their
ratio.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pyib_hip  # noqa: E402,F401
import pyib_hip.case as HC  # noqa: E402
import pyib_hip.mesh as PM  # noqa: E402
from pyib_hip.core import HipCore  # noqa: E402
from pyib_hip.driver import HipDriver  # noqa: E402
from oracle.pyib_ref import case as OC  # noqa: E402
from oracle.pyib_ref import core as OCore  # noqa: E402
from oracle.pyib_ref import driver as OD  # noqa: E402
from tests.helpers import (force_history, force_rel, golden, golden_config, parity_report, parity_tol,  # noqa: E402
                           progress, rel, small_case)


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def mesh3():
    return PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], 3))


@pytest.fixture(scope="module")
def cores(mesh3):
    print("[gpu] building HipCore on a 3-level nested box", flush=True)
    return HipCore(mesh3), OCore.Core(mesh3)


def test_incidence_operators(cores):
    hc, oc = cores
    rng = np.random.default_rng(0)
    e, f = rng.standard_normal(oc.nedge), rng.standard_normal(oc.nface)
    v = rng.standard_normal((oc.ncell, 3))
    assert rel(H(hc.curl(T(e))), oc.curl(e)) < 1e-15
    assert rel(H(hc.rot(T(f))), oc.rot(f)) < 1e-15
    assert rel(H(hc.interp_f2c(T(f))), oc.interp_f2c(f)) < 1e-14
    assert rel(H(hc.integr_c2f(T(v))), oc.integr_c2f(v)) < 1e-14
    assert rel(H(hc.cdiv(T(f))), oc.cdiv(f)) < 1e-14


@pytest.mark.parametrize("alpd", [0.0, 0.5])
def test_flow_mat(cores, alpd):
    hc, oc = cores
    rng = np.random.default_rng(1)
    p = rng.standard_normal(oc.nedge)
    kw = dict(alpd=alpd, visc=0.45, rou=1.0, dt=0.0043633)
    assert rel(H(hc.flow_mat(T(p), **kw)), oc.flow_mat(p.copy(), **kw)) < 1e-13


def _drivers(mesh, cavity, visc=0.05, check=20, structured=True):
    hc, oc = HipCore(mesh), OCore.Core(mesh)
    hd = HipDriver(hc, mesh, visc=visc, cg_max_iterations=3000, cg_check_interval=check, structured=structured)
    od = OD.Driver(oc, mesh, visc=visc, cg_max_iterations=3000, cg_check_interval=check)
    if cavity:
        hd.set_cavity(U_lid=0.0, spanwise_zerograd=False)
        od.set_cavity(U_lid=0.0, spanwise_zerograd=False)
    return hd, od


@pytest.mark.parametrize("cavity", [False, True])
def test_calc_vel_conv_diff(mesh3, cavity):
    hd, od = _drivers(mesh3, cavity)
    rng = np.random.default_rng(2)
    s = rng.standard_normal(od.core.nedge)
    s_h, U_h, vc_h = hd.calc_vel(T(s))
    s_o, U_o, vc_o = od.calc_vel(s)
    assert rel(H(s_h), s_o) == 0.0
    assert rel(H(U_h), U_o) < 1e-15
    assert rel(H(vc_h), vc_o) < 1e-14
    conv_h, diff_h = hd.conv_diff(U_h, vc_h)
    conv_o, diff_o = od.conv_diff(U_o, vc_o)
    assert rel(H(conv_h), conv_o) < 1e-13
    assert rel(H(diff_h), diff_o) < 1e-13


@pytest.mark.parametrize("cavity,structured", [(False, True), (True, True), (False, False), (True, False)])
def test_helm_solve(mesh3, cavity, structured):
    hd, od = _drivers(mesh3, cavity, visc=0.45, structured=structured)
    dt = 0.0043633
    rng = np.random.default_rng(3)
    s = rng.standard_normal(od.core.nedge) * 1e-2
    # manufactured right-hand side in range(A) (A has a gauge null space off the Dirichlet edges)
    xt = rng.standard_normal(od.core.nedge)
    res = od.apply_A(xt, dt)
    res[od.dir_mask] = 0.0
    x_o, m_o = od.helm_solve(res.copy(), s.copy(), dt)
    od.enable_sparse(dt)
    x_c, m_c = od.helm_solve(res.copy(), s.copy(), dt)
    floor = rel(x_c, x_o)
    x_h, m_h = hd.helm_solve(T(res), T(s), dt)
    progress(f"[gpu] helm_solve cavity={cavity} structured={structured}: iterations hip={m_h} oracle={m_o} "
             f"csr={m_c}")
    assert m_h == m_o == m_c
    assert parity_report(f"helm_solve cavity={cavity} structured={structured} solution vs oracle", rel(H(x_h), x_o),
                         floor)
    # the converged solve's true-residual ratio sits at round-off level; its spread between the
    # reference's own two paths sets the bound
    r_csr = od.last_helm_tolerance_ratio
    od.Asp = None
    od.helm_solve(res.copy(), s.copy(), dt)
    r_mf = od.last_helm_tolerance_ratio
    print(f"[gpu] tolerance ratio hip={hd.last_helm_tolerance_ratio:.6f} oracle={r_mf:.6f} csr={r_csr:.6f}")
    assert hd.last_helm_tolerance_ratio < 1.0 and hd.last_helm_converged
    assert abs(hd.last_helm_tolerance_ratio - r_mf) <= max(1e-6 * r_mf, 10.0 * abs(r_csr - r_mf))


@pytest.mark.parametrize("cavity", [False, True])
def test_grid_pcg_preconditioner_classes_change_no_bit(mesh3, cavity):
    """The PCG iteration with the preconditioner classes (prec from the per-level table, empty
    slots and Dirichlet rows skipped) against the same iterations reading the whole prec array:
    x, r, p and every CG scalar equal bit for bit, in float64 and float32, across the plain and
    the skip-last iteration blocks."""
    import ctypes as C
    from pyib_hip import _native as N
    from pyib_hip.core import dev_stream, ptr
    hc = HipCore(mesh3)
    hd = HipDriver(hc, mesh3, visc=0.05, cg_max_iterations=3000, structured=True, precision="mixed")
    if cavity:
        hd.set_cavity(U_lid=0.0, spanwise_zerograd=False)
    g = hd.grid
    for prec in ("float64", "float32"):
        st = g._struct[prec]
        assert st.edge_cls is not None and st.cls_prec is not None
        assert all(st.prec_level_g3[i] != 0.0 for i in range(st.nlevels))
    rng = np.random.default_rng(11)
    res = rng.standard_normal(hc.nedge)
    res[hd.dir_mask] = 0.0
    s0 = rng.standard_normal(hc.nedge) * 1e-2
    hdt = 0.5 * 0.0043633
    for prec, dt, pr in (("float64", torch.float64, hd.prec_i), ("float32", torch.float32, hd.prec_i32)):
        out = []
        for use_cls in (True, False):
            vec = {k: torch.zeros(g.nedge, dtype=dt, device="cuda") for k in "xrpw"}
            pr_use = pr if use_cls else pr.clone()
            scal = torch.zeros(N.NSCAL, dtype=torch.float64, device="cuda")
            cg = N.CG(ptr(vec["x"]), ptr(vec["r"]), ptr(vec["p"]), ptr(vec["w"]), ptr(pr_use), ptr(scal))
            r_i, s_i = g.to_internal(T(res), dtype=dt), g.to_internal(T(s0), dtype=dt)
            m, ws = g.struct(prec), g.work_struct(prec)
            N.call("pyib_grid_cg_begin", m, ws, C.byref(cg), ptr(r_i), ptr(s_i), 1e-10, 1e-9, dev_stream())
            N.call("pyib_grid_cg_iterate", m, ws, C.byref(cg), 1.0, 0.05, hdt, 37, 0, dev_stream())
            N.call("pyib_grid_cg_iterate", m, ws, C.byref(cg), 1.0, 0.05, hdt, 3, 1, dev_stream())
            N.call("pyib_grid_cg_direction", m, C.byref(cg), dev_stream())
            N.call("pyib_grid_cg_iterate", m, ws, C.byref(cg), 1.0, 0.05, hdt, 5, 0, dev_stream())
            torch.cuda.synchronize()
            out.append(([H(vec[k]) for k in "xrp"], H(scal)))
        for a, b in zip(out[0][0], out[1][0]):
            assert np.isfinite(a).all() and np.array_equal(a, b)
        assert np.array_equal(out[0][1], out[1][1])
        print(f"[gpu] prec classes cavity={cavity} {prec}: 45 iterations bitwise equal, eta {out[0][1][0]:.6e}")


@pytest.mark.parametrize("h,boxes", [(0.0625, "nested3"), (0.5 / 24, "nested2"), (0.125, "wake")])
def test_grid_flow_mat(h, boxes):
    """Structured stencils + composite rows == oracle flow_mat through the signed renumbering."""
    from pyib_hip.grid import HipGrid
    bx = {"nested3": PM.nested_box_boxes([1.5, 1.25, 1.25], 3), "nested2": PM.nested_box_boxes([1.0, 0.75, 0.75], 2),
          "wake": PM.wake_boxes()}[boxes]
    m = PM.make_nested_mesh(h, bx)
    hc, oc = HipCore(m), OCore.Core(m)
    g = HipGrid(hc, precisions=("float64", "float32"))
    g.set_dir_mask(np.zeros(oc.nedge, bool))
    rng = np.random.default_rng(8)
    p = rng.standard_normal(oc.nedge)
    kw = dict(visc=0.45, rou=1.0, dt=0.0043633)
    ref = oc.flow_mat(p.copy(), alpd=0.5, **kw)
    p_int = g.to_internal(T(p))
    w_int, dot = g.flow_mat(p_int, kw["rou"], kw["visc"], kw["dt"], with_dot=True)
    w = g.to_reference(w_int)
    assert rel(H(w), ref) < 1e-13
    # the PCG's p.A p, formed from the cell passes, against p.w and the oracle's p.(A p)
    pap = float(p @ ref)
    assert abs(float(dot.item()) - pap) <= 1e-12 * abs(pap)
    assert abs(float(dot.item()) - float((p_int * w_int).sum())) <= 1e-12 * abs(pap)
    p32 = g.to_internal(T(p), dtype=torch.float32)
    w32i, dot32 = g.flow_mat(p32, kw["rou"], kw["visc"], kw["dt"], with_dot=True)
    w32 = g.to_reference(w32i, dtype=torch.float64)
    assert rel(H(w32), ref) < 1e-5
    assert abs(float(dot32.item()) - pap) <= 1e-5 * abs(pap)


def test_grid_flow_mat_without_dirichlet_composite_rows():
    """With a cavity's Dirichlet mask the G3 composite drops its Dirichlet rows: w is still the
    oracle's flow_mat on every other edge, stays exactly 0 on the Dirichlet edges, and the
    PCG's p.A p (p zero there, as every direction is) is unchanged."""
    from pyib_hip.grid import HipGrid
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.25], 3))
    hc, oc = HipCore(m), OCore.Core(m)
    g = HipGrid(hc, precisions=("float64", "float32"))
    dm = np.asarray(m.bc_edge) > 0
    g.set_dir_mask(np.zeros(oc.nedge, bool))
    n_all = g._comp["g3"]["n"]
    g.set_dir_mask(dm)
    assert g._comp["g3"]["n"] < n_all
    rng = np.random.default_rng(9)
    p = rng.standard_normal(oc.nedge)
    p[dm] = 0.0
    kw = dict(visc=0.45, rou=1.0, dt=0.0043633)
    ref = oc.flow_mat(p.copy(), alpd=0.5, **kw)
    ref[dm] = 0.0
    w_int, dot = g.flow_mat(g.to_internal(T(p)), kw["rou"], kw["visc"], kw["dt"], with_dot=True)
    w = H(g.to_reference(w_int))
    assert rel(w, ref) < 1e-13
    assert np.all(w[dm] == 0.0)
    pap = float(p @ ref)
    assert abs(float(dot.item()) - pap) <= 1e-12 * abs(pap)
    w32 = H(g.to_reference(g.flow_mat(g.to_internal(T(p), dtype=torch.float32), kw["rou"], kw["visc"], kw["dt"]),
                           dtype=torch.float64))
    assert rel(w32, ref) < 1e-5
    assert np.all(w32[dm] == 0.0)


def _run_hip(cfg, steps):
    rt = HC.build_runtime(cfg)
    s, s_old, vc = HC.initial_state(rt)
    rows = []
    for step in range(1, steps + 1):
        rt.driver.time = (step - 1) * rt.dt
        s, vc, s_old = rt.driver.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        rt.driver.time = step * rt.dt
        rt.load_correction.advance(vc, HC.sphere_center(rt.config, rt.driver.time), rt.dt, rt.driver.ib_force_E)
        rows.append(HC.diagnostics(rt, step, s, vc, 0.0))
    return rt, s, vc, rows


@pytest.mark.parametrize("name,steps", [("cpu_fixed_sphere_smoke", 5), ("cpu_oscillating_sphere_smoke", 4),
                                        ("cpu_fixed_sphere_re100_d8_gate", 3)])
def test_reference_golden_trajectory(name, steps):
    """Same case, same steps as the reference's own CPU run: identical CG counts, s to 1e-12."""
    cfg = golden_config(name)
    cfg["runtime"]["device"] = "hip"
    cfg = HC.validate_config(cfg)
    print(f"[gpu] golden case {name}", flush=True)
    rt, s, vc, rows = _run_hip(cfg, steps)
    ort = OC.build_runtime(cfg, sparse=False)
    os_mf = OC.run(ort, steps)[0]
    fl_mf = np.asarray(ort.driver.ib_force_L)
    with golden(f"traj_{name}.npz") as g:
        cols = [str(c) for c in g["columns"]]
        hist = g["history"]
        it = cols.index("helmholtz_iterations")
        assert [r["helmholtz_iterations"] for r in rows] == [int(v) for v in hist[:, it]]
        floor = rel(os_mf, g["s"])        # the matrix-free oracle against the reference's own run
        assert parity_report(f"{name}: s vs reference", rel(H(s), g["s"]), floor)
        ffl = rel(fl_mf, g["ib_force_L"])
        print(f"[gpu] {name}: Lagrangian force vs reference {rel(H(rt.driver.ib_force_L), g['ib_force_L']):.2e} "
              f"(matrix-free oracle vs reference {ffl:.2e})")
        assert rel(H(rt.driver.ib_force_L), g["ib_force_L"]) <= parity_tol(ffl, base=1e-10)
        fr = force_rel(force_history(rows), force_history((hist, cols)))
        fc = force_rel(force_history(rows, ("force_fluid_x", "Cd")), force_history((hist, cols), ("force_fluid_x", "Cd")))
        print(f"[gpu] {name}: force history rel {fr:.2e}, (force_fluid_x, Cd) rel {fc:.2e}")
        assert fr <= 1e-10 and fc <= 1e-10


@pytest.mark.parametrize("kind", ["oscillating", "fixed"])
def test_nested_box_steps_vs_oracle(kind):
    cfg = small_case(kind, steps=3)
    print(f"[gpu] nested_box {kind} 3 steps", flush=True)
    rt, s, vc, rows = _run_hip(cfg, 3)
    ort = OC.build_runtime(cfg, sparse=False)
    os_, oso, ovc, orows = OC.run(ort, 3)
    crt = OC.build_runtime(cfg, sparse=True)
    cs, _, cvc, crows = OC.run(crt, 3)
    assert [r["helmholtz_iterations"] for r in rows] == [r["helmholtz_iterations"] for r in orows] == \
        [r["helmholtz_iterations"] for r in crows]
    vort = lambda core, v: core.rot(core.integr_c2f(v[:core.ncell]))
    om, cm = vort(ort.core, ovc), vort(crt.core, cvc)
    hm = H(rt.core.rot(rt.core.integr_c2f(vc[:rt.core.ncell].contiguous())))
    fs, fw = rel(cs, os_), rel(cm, om)
    assert parity_report(f"nested_box {kind}: s vs oracle", rel(H(s), os_), fs)
    assert parity_report(f"nested_box {kind}: vorticity vs oracle", rel(hm, om), fw)
    fr = force_rel(force_history(rows), force_history(orows))
    ff = force_rel(force_history(crows), force_history(orows))
    print(f"[gpu] nested_box {kind}: force history rel {fr:.2e} (floor {ff:.2e})")
    assert fr <= max(1e-10, 10.0 * ff)


def test_bitwise_reproducible():
    cfg = small_case("oscillating", steps=2)
    a = H(_run_hip(cfg, 2)[1])
    b = H(_run_hip(cfg, 2)[1])
    assert np.array_equal(a, b)
