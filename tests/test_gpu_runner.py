"""Run-level behaviour on the GPU: IB transfer, loads, resume, strict CG, mixed precision.

Tolerances: IB interpolation/spreading 1e-15 relative vs the oracle and
bitwise run-to-run; load reductions 1e-13; resume bitwise identical to an
uninterrupted run; mixed precision vs float64 <= 1e-6 relative on the forces.
"""
import copy
import json
import subprocess
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pyib_hip  # noqa: E402,F401
import pyib_hip.case as HC  # noqa: E402
import pyib_hip.mesh as PM  # noqa: E402
from pyib_hip.immersed import HipImmersedBoundary  # noqa: E402
from pyib_hip.interior import HipLoadCorrection  # noqa: E402
from oracle.pyib_ref import immersed as OI  # noqa: E402
from oracle.pyib_ref import interior as OIn  # noqa: E402
from oracle.pyib_ref import core as OCore  # noqa: E402
from tests.helpers import ROOT, rel, small_case  # noqa: E402


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_ib_transfer_matches_oracle_and_is_deterministic():
    mesh = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.0, 1.0, 1.0], 2))
    oc = OCore.Core(mesh)
    rng = np.random.default_rng(5)
    mk = rng.uniform(-0.4, 0.4, size=(300, 3))
    ds = rng.uniform(0.03, 0.07, 300)
    f = rng.standard_normal((300, 3))
    vc = rng.standard_normal((oc.ncell, 3))
    oib = OI.ImmersedBoundary(mesh)
    hib = HipImmersedBoundary(mesh)
    hib.stencil(T(mk))
    vel = hib.interp(T(vc)).cpu().numpy()
    assert rel(vel, oib.interp(vc, mk)) < 1e-15
    out1 = torch.zeros((oc.ncell, 3), dtype=torch.float64, device="cuda")
    hib.spread_into(out1, T(f), T(ds ** 3))
    ref = oib.spread(f, mk, ds)
    assert rel(out1.cpu().numpy(), ref) < 1e-15
    out2 = torch.zeros_like(out1)
    hib.stencil(T(mk))
    hib.spread_into(out2, T(f), T(ds ** 3))
    assert torch.equal(out1, out2)
    # accumulate mode: out += 2*spread
    hib.spread_into(out2, T(f), T(ds ** 3), scale=2.0, accumulate=True)
    assert rel(out2.cpu().numpy(), 3.0 * ref) < 1e-15


def test_spread_runs_longer_than_a_block():
    """Markers packed into one cell give every touched cell a run of 400 taps, longer than the
    256 taps one block reduces in LDS: the segment head finishes the run across block
    boundaries, in marker order, and the result still equals the oracle's bincount."""
    mesh = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.0, 1.0, 1.0], 2))
    oc = OCore.Core(mesh)
    rng = np.random.default_rng(7)
    n = 400
    mk = 0.01 * rng.uniform(-1.0, 1.0, size=(n, 3))
    ds = rng.uniform(0.03, 0.07, n)
    f = rng.standard_normal((n, 3))
    oib = OI.ImmersedBoundary(mesh)
    hib = HipImmersedBoundary(mesh)
    hib.stencil(T(mk))
    out = torch.zeros((oc.ncell, 3), dtype=torch.float64, device="cuda")
    hib.spread_into(out, T(f), T(ds ** 3))
    ref = oib.spread(f, mk, ds)
    got = out.cpu().numpy()
    print(f"[gpu] long-run spreading: rel {rel(got, ref):.3e}, bitwise {np.array_equal(got, ref)}")
    assert rel(got, ref) < 1e-15
    again = torch.zeros_like(out)
    hib.spread_into(again, T(f), T(ds ** 3))
    assert torch.equal(out, again)


def test_load_reductions_match_oracle():
    mesh = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.0, 1.0, 1.0], 2))
    oc = OCore.Core(mesh)
    rng = np.random.default_rng(6)
    pos, vol = oc.cpos, 1.0 / oc.vol_ci
    kw = dict(reference_center=[0, 0, 0], maximum_translation=[0.05, 0, 0], radius=0.25, smoothing_width=0.0625,
              density=1.3, force_support_width=1.5 * 0.0625)
    o = OIn.ClosedSphereLoadCorrection(pos, vol, **kw)
    h = HipLoadCorrection(pos, vol, **kw)
    v0, v1 = rng.standard_normal((oc.ncell, 3)), rng.standard_normal((oc.ncell, 3))
    F = rng.standard_normal((oc.ncell, 3))
    o.initialize(v0, [0.01, 0, 0])
    h.initialize(T(v0), [0.01, 0, 0])
    lo = o.advance(v1, [0.02, 0, 0], 0.01, F)
    lh = h.advance(T(v1), [0.02, 0, 0], 0.01, T(F))
    for name in ("force_fluid", "moment_fluid", "body_force", "body_moment"):
        assert rel(getattr(lh, name), getattr(lo, name)) < 1e-12, name
    assert abs(float(lh.interior_momenta.indicator_volume) - float(lo.interior_momenta.indicator_volume)) < 1e-13


def _cfg(tmp_steps, **over):
    cfg = small_case("oscillating", steps=tmp_steps)
    cfg["output"] = {"sample_every": 1, "checkpoint_every": 2, "field_every": 0, "field_start_step": 0,
                     "save_field": True}
    for k, v in over.items():
        cfg[k].update(v)
    return HC.validate_config(cfg)


def test_resume_is_bitwise_identical(tmp_path):
    full = HC.run_case(_cfg(4), tmp_path / "a", quiet=True)
    part = HC.run_case(_cfg(2), tmp_path / "b", quiet=True)
    assert part.completed_step == 2
    res = HC.run_case(_cfg(4), tmp_path / "b", resume=tmp_path / "b" / "checkpoint.npz", quiet=True)
    assert res.start_step == 2 and res.completed_step == 4
    assert np.array_equal(full.s, res.s) and np.array_equal(full.ib_force_E, res.ib_force_E)
    ha = [{k: v for k, v in r.items() if k != "milliseconds_per_step"} for r in full.history]
    hb = [{k: v for k, v in r.items() if k != "milliseconds_per_step"} for r in res.history]
    assert ha == hb
    with np.load(tmp_path / "a" / "final_field.npz", allow_pickle=False) as ff:
        assert str(ff["device"]) == "hip" and ff["vc"].shape[1] == 3
    md = json.loads((tmp_path / "a" / "metadata.json").read_text())
    assert md["complete"] and md["backend"] == "hip/pyib_hip"


def test_resume_rejects_foreign_checkpoint(tmp_path):
    HC.run_case(_cfg(2), tmp_path / "a", quiet=True)
    other = _cfg(4, physics={"reynolds": 0.3})
    with pytest.raises(HC.CaseConfigError, match="configuration"):
        HC.run_case(other, tmp_path / "a", resume=tmp_path / "a" / "checkpoint.npz", quiet=True)


def test_strict_cg_failure_exit_code(tmp_path):
    raw = copy.deepcopy(small_case("oscillating", steps=1))
    raw["solver"]["cg_max_iterations"] = 3
    raw["solver"]["fail_on_cg_nonconvergence"] = True
    case = tmp_path / "c.json"
    case.write_text(json.dumps(raw))
    r = subprocess.run([sys.executable, "-m", "pyib_hip.cli", str(case), "--outdir", str(tmp_path / "o"), "--quiet"],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "did not converge" in r.stderr


def _forces(cfg, steps):
    rt = HC.build_runtime(cfg)
    s, s_old, vc = HC.initial_state(rt)
    out = []
    for step in range(1, steps + 1):
        rt.driver.time = (step - 1) * rt.dt
        s, vc, s_old = rt.driver.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        rt.driver.time = step * rt.dt
        out.append(np.asarray(rt.load_correction.advance(vc, HC.sphere_center(rt.config, rt.driver.time), rt.dt,
                                                         rt.driver.ib_force_E).body_force))
    return np.array(out), s.cpu().numpy(), rt


def test_mixed_precision_force_error():
    """Mixed precision vs float64: force history to 1e-6 relative; the face-flux field error is
    stated and bounded by 1e-5 (float32 rounding in the operator's slowest modes, which the
    residual stop test does not resolve; see test_mixed_precision_field_error_sources).  The
    stream function itself also carries a gauge (gradient) component that the equations leave
    free; it is printed for reference only."""
    f64, s64, r64 = _forces(small_case("oscillating", steps=4), 4)
    fmx, smx, rmx = _forces(small_case("oscillating", precision="mixed", steps=4), 4)
    err = np.abs(fmx - f64).max() / np.abs(f64).max()
    U64 = r64.core.curl(torch.from_numpy(s64).cuda()).cpu().numpy()
    Umx = rmx.core.curl(torch.from_numpy(smx).cuda()).cpu().numpy()
    print(f"[gpu] mixed vs float64: force rel err {err:.2e}, face flux rel err {rel(Umx, U64):.2e}, "
          f"stream function (incl. gauge) {rel(smx, s64):.2e}")
    assert err <= 1e-6
    assert rel(Umx, U64) <= 1e-5
    assert rmx.driver.helmholtz_converged_last_step


def test_mixed_precision_field_error_sources():
    """Where mixed precision and float64 differ, against a float64 solve 100x tighter (1e-12 /
    1e-11) than the reference's tolerance (1e-10 / 1e-9).  The edge vorticity rot(integr(vc)) is
    the operator applied to s, which the stop test bounds: both precisions lie within a few times
    the tolerance of the tight solve.  The face flux also carries the components the residual test
    cannot see (the operator's slowest modes, whose residual is small for a visible error): float64
    at the reference's tolerance stays within 1e-7 of the tight solve, while mixed carries float32
    rounding there, stated and bounded by 1e-5."""
    import functools

    def run(precision, tight=False):
        rt = HC.build_runtime(small_case("oscillating", precision=precision, steps=4))
        if tight:
            rt.driver.helm_solve = functools.partial(rt.driver.helm_solve, tol1=1e-12, tol2=1e-11)
        s, s_old, vc = HC.initial_state(rt)
        for step in range(1, 5):
            rt.driver.time = (step - 1) * rt.dt
            s, vc, s_old = rt.driver.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
            assert rt.driver.helmholtz_converged_last_step
        U = rt.core.curl(s).cpu().numpy()
        W = rt.core.rot(rt.core.integr_c2f(vc[:rt.core.ncell].contiguous())).cpu().numpy()
        return U, W

    Ut, Wt = run("float64", tight=True)
    U64, W64 = run("float64")
    Umx, Wmx = run("mixed")
    w64, wmx = rel(W64, Wt), rel(Wmx, Wt)
    u64, umx = rel(U64, Ut), rel(Umx, Ut)
    print(f"[gpu] against float64 at 1e-12/1e-11: edge vorticity float64 {w64:.2e}, mixed {wmx:.2e}; "
          f"face flux float64 {u64:.2e}, mixed {umx:.2e}")
    assert w64 <= 1e-9 and wmx <= 1e-9
    assert u64 <= 1e-7
    assert umx <= 1e-5


def test_reference_d32_run_history():
    """The reference's own D32 oscillating-sphere run (5-level concentric mesh, CUDA FP64,
    checks/pressure-validation-20260925.zip runs/D32): the first sampled steps of its
    history.csv against the HIP path on the same case file."""
    import csv
    import os
    from tests.helpers import GOLDEN
    ref = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))["D32"]["case"]
    rows = list(csv.DictReader(open(os.path.join(GOLDEN, "reference_d32_history_head.csv"))))
    want = [r for r in rows if int(r["step"]) <= 10]
    cfg = copy.deepcopy(ref)
    cfg["runtime"]["device"] = "hip"
    cfg["time"]["steps"] = 10
    cfg = HC.validate_config(cfg)
    print("[gpu] reference D32 case, 10 steps", flush=True)
    rt = HC.build_runtime(cfg)
    s, s_old, vc = HC.initial_state(rt)
    got = {}
    for step in range(1, 11):
        rt.driver.time = (step - 1) * rt.dt
        s, vc, s_old = rt.driver.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        rt.driver.time = step * rt.dt
        rt.load_correction.advance(vc, HC.sphere_center(rt.config, rt.driver.time), rt.dt, rt.driver.ib_force_E)
        got[step] = HC.diagnostics(rt, step, s, vc, 0.0)
    worst = {}
    for r in want:
        g = got[int(r["step"])]
        print(f"[gpu] step {r['step']}: CG {g['helmholtz_iterations']} vs {r['helmholtz_iterations']}, "
              f"Fx {g['body_force_x']:.15e} vs {float(r['body_force_x']):.15e}")
        assert g["helmholtz_iterations"] == int(r["helmholtz_iterations"])
        for col in ("body_force_x", "force_fluid_x", "interior_momentum_x", "slip_max", "speed_max"):
            a, b = g[col], float(r[col])
            worst[col] = max(worst.get(col, 0.0), abs(a - b) / max(abs(b), 1e-300))
    print("[gpu] worst relative differences vs the reference's CUDA run:", worst)
    assert worst["body_force_x"] < 1e-8 and worst["force_fluid_x"] < 1e-8


def test_mixed_precision_force_harmonics_over_a_period():
    """Mixed vs float64 over one full period (D8 nested box, dt = T/100): the harmonic vector
    (a cos + b sin + c fit of body_force_x over the period, the reference's validation fit
    tools/check_pressure_validation.py:32-38) agrees to 1e-6 relative."""
    from oracle.pyib_ref.case import harmonic_fit
    f = 0.5729577951308232
    steps = 100
    over = {"time": {"dt": (1.0 / f) / steps, "steps": steps, "inner_iterations": 4}}
    f64, _, _ = _forces(small_case("oscillating", steps=steps, **over), steps)
    fmx, _, _ = _forces(small_case("oscillating", precision="mixed", steps=steps, **over), steps)
    t = (1.0 / f) / steps * np.arange(1, steps + 1)
    h64 = harmonic_fit(t, f64[:, 0], 2.0 * np.pi * f, 0.0)
    hmx = harmonic_fit(t, fmx[:, 0], 2.0 * np.pi * f, 0.0)
    err = float(np.linalg.norm(hmx - h64) / np.linalg.norm(h64))
    print(f"[gpu] one period, mixed vs float64: harmonic (a, b, c) = {h64} vs {hmx}, rel err {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("name", ["cpu_fixed_sphere_smoke", "cpu_oscillating_sphere_smoke"])
def test_run_case_history_matches_reference_and_validates(tmp_path, name):
    """run_case on the reference's bundled smoke case (its own case file, device hip, final field
    saved): the history.csv it writes matches the reference's own run column by column
    (tests.helpers.compare_history), and the artifact validator accepts the directory;
    tampered copies are refused with the reason."""
    import shutil
    import pyib_hip.validate as V
    from tests.helpers import compare_history, golden, golden_config
    cfg = golden_config(name)
    cfg["runtime"]["device"] = "hip"
    cfg["solver"]["sparse"] = False
    cfg["output"]["save_field"] = True
    cfg = HC.validate_config(cfg)
    out = tmp_path / "run"
    HC.run_case(cfg, out, quiet=True)
    rows = HC.read_history(out / "history.csv")
    with golden(f"traj_{name}.npz") as g:
        report = compare_history(rows, g["history"], [str(c) for c in g["columns"]])
    print(f"[gpu] {name} run_case history.csv vs the reference's, worst / allowance:",
          {k: round(v, 3) for k, v in sorted(report.items(), key=lambda kv: -kv[1])[:5]})
    summary = V.validate_result(out)
    assert summary["coverage"] == "full_field" and summary["completed_step"] == cfg["time"]["steps"]
    assert V.main([str(out)]) == 0

    def tampered(tag, edit):
        d = tmp_path / tag
        shutil.copytree(out, d)
        edit(d)
        with pytest.raises(V.ValidationError) as e:
            V.validate_result(d)
        return str(e.value)

    def device(d):
        with np.load(d / "checkpoint.npz", allow_pickle=False) as a:
            arr = {k: a[k] for k in a.files}
        arr["device"] = np.str_("cuda")
        np.savez(d / "checkpoint.npz", **arr)

    def sources(d):
        md = json.loads((d / "metadata.json").read_text())
        md["source_sha256"]["driver.py"] = "0" * 64
        (d / "metadata.json").write_text(json.dumps(md))

    def history(d):
        lines = (d / "history.csv").read_text().splitlines()
        (d / "history.csv").write_text("\n".join(lines[:-2] + lines[-1:]) + "\n")

    def field(d):
        with np.load(d / "final_field.npz", allow_pickle=False) as a:
            arr = {k: a[k] for k in a.files}
        arr["ib_force_E"] = arr["ib_force_E"] * (1.0 + 1e-6)
        np.savez(d / "final_field.npz", **arr)

    assert "device" in tampered("device", device)
    assert "source hash" in tampered("sources", sources)
    assert "omits required sampled steps" in tampered("history", history)
    assert "differs between checkpoint and final field" in tampered("field", field)
