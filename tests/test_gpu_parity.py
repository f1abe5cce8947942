"""HIP path vs the CPU oracle and the reference's golden runs (needs a GPU).

Tolerances (float64 path): incidence operators 1e-14 relative L2, elliptic
operator 1e-13, Helmholtz CG identical iteration count and 1e-12 on the
solution, full steps 1e-12 on s / edge circulation and 1e-10 on forces.
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
from tests.helpers import golden, golden_config, rel, small_case  # noqa: E402


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
    x_h, m_h = hd.helm_solve(T(res), T(s), dt)
    print(f"[gpu] helm_solve cavity={cavity} structured={structured}: iterations hip={m_h} oracle={m_o}, "
          f"rel={rel(H(x_h), x_o):.2e}")
    assert m_h == m_o
    assert rel(H(x_h), x_o) < 1e-12
    assert abs(hd.last_helm_tolerance_ratio - od.last_helm_tolerance_ratio) < 1e-6 * od.last_helm_tolerance_ratio


@pytest.mark.parametrize("h,boxes", [(0.0625, "nested3"), (0.5 / 24, "nested2"), (0.125, "wake")])
def test_grid_flow_mat(h, boxes):
    """Structured stencils + composite rows == oracle flow_mat through the signed renumbering."""
    from pyib_hip.grid import HipGrid
    bx = {"nested3": PM.nested_box_boxes([1.5, 1.25, 1.25], 3), "nested2": PM.nested_box_boxes([1.0, 0.75, 0.75], 2),
          "wake": PM.wake_boxes()}[boxes]
    m = PM.make_nested_mesh(h, bx)
    hc, oc = HipCore(m), OCore.Core(m)
    g = HipGrid(hc, precisions=("float64", "float32"))
    rng = np.random.default_rng(8)
    p = rng.standard_normal(oc.nedge)
    kw = dict(visc=0.45, rou=1.0, dt=0.0043633)
    ref = oc.flow_mat(p.copy(), alpd=0.5, **kw)
    w = g.to_reference(g.flow_mat(g.to_internal(T(p)), kw["rou"], kw["visc"], kw["dt"]))
    assert rel(H(w), ref) < 1e-13
    w32 = g.to_reference(g.flow_mat(g.to_internal(T(p), dtype=torch.float32), kw["rou"], kw["visc"], kw["dt"]),
                         dtype=torch.float64)
    assert rel(H(w32), ref) < 1e-5


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
    with golden(f"traj_{name}.npz") as g:
        cols = [str(c) for c in g["columns"]]
        hist = g["history"]
        it = cols.index("helmholtz_iterations")
        assert [r["helmholtz_iterations"] for r in rows] == [int(v) for v in hist[:, it]]
        assert rel(H(s), g["s"]) < 1e-12
        assert rel(H(rt.driver.ib_force_L), g["ib_force_L"]) < 1e-10
        for col in ("body_force_x", "body_force_y", "body_force_z", "force_fluid_x", "Cd"):
            k = cols.index(col)
            ours = np.array([r[col] for r in rows])
            assert np.abs(ours - hist[:, k]).max() <= 1e-10 * np.abs(hist[:, k]).max() + 1e-15, col


@pytest.mark.parametrize("kind", ["oscillating", "fixed"])
def test_nested_box_steps_vs_oracle(kind):
    cfg = small_case(kind, steps=3)
    print(f"[gpu] nested_box {kind} 3 steps", flush=True)
    rt, s, vc, rows = _run_hip(cfg, 3)
    ort = OC.build_runtime(cfg, sparse=False)
    os_, oso, ovc, orows = OC.run(ort, 3)
    assert [r["helmholtz_iterations"] for r in rows] == [r["helmholtz_iterations"] for r in orows]
    assert rel(H(s), os_) < 1e-12
    om = ort.core.rot(ort.core.integr_c2f(ovc[:ort.core.ncell]))
    hm = H(rt.core.rot(rt.core.integr_c2f(vc[:rt.core.ncell].contiguous())))
    assert rel(hm, om) < 1e-12
    for col in ("body_force_x", "body_force_y", "body_force_z"):
        ours = np.array([r[col] for r in rows])
        ref = np.array([r[col] for r in orows])
        assert np.abs(ours - ref).max() <= 1e-10 * np.abs(ref).max() + 1e-15, col


def test_bitwise_reproducible():
    cfg = small_case("oscillating", steps=2)
    a = H(_run_hip(cfg, 2)[1])
    b = H(_run_hip(cfg, 2)[1])
    assert np.array_equal(a, b)
