"""Implicit IB forcing (ib_immersed.py:122-174) and the rotating-sphere case (ib_case.py:566-584)
on the GPU against the oracle.

* The implicit Lagrangian force f = M^-1 (u_d - interp)/dt and its spread field equal the
  oracle's to 1e-11 relative (the coupling is inverted on each side; M is well conditioned at
  marker spacing h), and they remove the slip: interp(vc + dt F_E) = u_d to 1e-6 of the slip.
* Whole runs, 2 steps each on a uniform 16^3 box (50 markers at spacing h), compared with the
  oracle's run of the same case by the per-column history rules (tests.helpers.compare_history):
  a fixed sphere with implicit direct forcing (alpha 0.5) and a rotating sphere (spin ratio 0.5)
  with explicit accumulated forcing.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pyib_hip  # noqa: E402,F401
import pyib_hip.case as HC  # noqa: E402
import pyib_hip.mesh as PM  # noqa: E402
from pyib_hip.immersed import HipImmersedBoundary  # noqa: E402
from oracle.pyib_ref import case as OC  # noqa: E402
from oracle.pyib_ref import immersed as OI  # noqa: E402
from tests.helpers import compare_history, rel, small_case  # noqa: E402


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_implicit_force_matches_oracle_and_removes_slip():
    mesh = PM.make_nested_mesh(0.0625, PM.uniform_boxes([[-0.5, 0.5], [-0.5, 0.5], [-0.5, 0.5]]))
    n = max(12, int(round(4 * np.pi * 0.25 ** 2 / 0.0625 ** 2)))
    mk = OC.fibonacci_sphere(n, 0.25, [0.01, -0.02, 0.015])
    ds = np.full(n, (4 * np.pi * 0.25 ** 2 * 0.0625 / n) ** (1 / 3))
    rng = np.random.default_rng(2)
    oib = OI.ImmersedBoundary(mesh)
    vc = rng.standard_normal((oib.ncell, 3))
    ud = rng.standard_normal((n, 3))
    dt = 0.01
    oE, of = oib.implicit_force(vc, mk, ud, dt, ds)
    hib = HipImmersedBoundary(mesh)
    hib.stencil(T(mk), version=1)
    hib.interp(T(vc))
    f = hib.implicit_lagrangian_force(T(ud), dt, T(ds ** 3))
    E = torch.zeros((oib.ncell, 3), dtype=torch.float64, device="cuda")
    hib.spread_into(E, f, T(ds ** 3))
    fh, Eh = f.cpu().numpy(), E.cpu().numpy()
    slip0 = np.linalg.norm(oib.interp(vc, mk) - ud)
    after = np.linalg.norm(hib.interp(T(vc) + dt * E).cpu().numpy() - ud)
    print(f"[gpu] implicit force: f vs oracle {rel(fh, of):.2e}, F_E {rel(Eh, oE):.2e}, slip {after / slip0:.2e}",
          flush=True)
    assert rel(fh, of) < 1e-11 and rel(Eh, oE) < 1e-11
    assert after <= 1e-6 * slip0
    # the cached inverse is reused for the same stencil and rebuilt for a moved one
    key = hib._imp_key
    hib.interp(T(vc))
    hib.implicit_lagrangian_force(T(ud), dt, T(ds ** 3))
    hib.stencil(T(mk + 0.003), version=2)
    hib.interp(T(vc))
    hib.implicit_lagrangian_force(T(ud), dt, T(ds ** 3))
    assert hib._imp_key != key


def _runs(cfg, steps):
    rt = HC.build_runtime(cfg)
    s, s_old, vc = HC.initial_state(rt)
    rows = []
    for step in range(1, steps + 1):
        rt.driver.time = (step - 1) * rt.dt
        s, vc, s_old = rt.driver.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        rt.driver.time = step * rt.dt
        rt.load_correction.advance(vc, HC.sphere_center(rt.config, rt.driver.time), rt.dt, rt.driver.ib_force_E)
        rows.append(HC.diagnostics(rt, step, s, vc, 0.0))
    ort = OC.build_runtime(cfg, sparse=False)
    orows = OC.run(ort, steps)[3]
    cols = [c for c in HC.HISTORY_COLUMNS if c != "milliseconds_per_step"]
    hist = np.array([[float(r[c]) for c in cols] for r in orows])
    return rows, hist, cols, rt


BOX16 = {"kind": "uniform_box", "dx": 0.125, "bounds": [[-1, 1], [-1, 1], [-1, 1]]}


@pytest.mark.parametrize("kind", ["implicit_fixed", "rotating"])
def test_case_runs_match_the_oracle(kind):
    if kind == "implicit_fixed":
        cfg = small_case("fixed", steps=2, mesh=BOX16,
                         ib={"force_mode": "direct", "implicit": True, "alpha": 0.5, "marker_spacing": 0.125})
    else:
        cfg = small_case("fixed", steps=2, mesh=BOX16, case={"kind": "rotating_sphere"},
                         ib={"marker_spacing": 0.125,
                             "motion": {"kind": "rotation", "axis": [0.0, 0.0, 1.0], "spin_ratio": 0.5}})
    assert cfg["ib"]["implicit"] == (kind == "implicit_fixed")
    print(f"[gpu] {kind}: 2 steps against the oracle", flush=True)
    rows, hist, cols, rt = _runs(cfg, 2)
    report = compare_history(rows, hist, cols)
    print(f"[gpu] {kind} history columns, worst / allowance:",
          {k: round(v, 4) for k, v in sorted(report.items(), key=lambda kv: -kv[1])[:6]}, flush=True)
    if kind == "rotating":
        v = rt.driver.ib_vel.cpu().numpy()
        assert np.abs(v[:, 2]).max() == 0.0 and np.abs(v).max() > 0.0
