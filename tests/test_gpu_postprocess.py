"""GPU pressure recovery (pyib_hip.postprocess) against the oracle's restatement of the reference
(oracle/pyib_ref/postprocess.py, ib_postprocess.py:49-105).

* The reference's own test of recover_pressure (tests/test_ib_postprocess.py:76-85): a quadratic
  pressure on its 9^3 graph is recovered from its exact gradient up to the boundary gauge (atol
  3e-9, rtol 1e-8), boundary mean below 1e-14, projection residual below 1e-9.
* A random gradient at a tight tolerance (1e-13 for both): the GPU pressure equals the oracle's
  and the dense solve to 1e-9 of its size.
* A run's own final_field.npz (run_case on the reference's bundled oscillating smoke case):
  the pressure of the CLI equals the oracle's on the same snapshot to 1e-8 of its size, at the
  reference's default tolerance 1e-10 (the two CGs stop a few iterations apart).
"""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pyib_hip  # noqa: E402,F401
import pyib_hip.case as HC  # noqa: E402
import pyib_hip.postprocess as PP  # noqa: E402
from oracle.pyib_ref import postprocess as OP  # noqa: E402
from tests.helpers import golden_config  # noqa: E402
from tests.test_oracle_postprocess import quadratic, quadratic_gradient, uniform_graph  # noqa: E402


def test_quadratic_pressure_is_recovered():
    pos, faces, area = uniform_graph(9)
    p, diag = PP.recover_pressure(pos, faces, area, quadratic_gradient(pos))
    want = quadratic(pos)
    gauge = np.unique(faces[faces[:, 1] < 0, 0])
    want -= want[gauge].mean()
    print(f"[gpu] quadratic pressure: max err {np.abs(p - want).max():.2e}, {diag}", flush=True)
    np.testing.assert_allclose(p, want, atol=3e-9, rtol=1e-8)
    assert abs(p[gauge].mean()) < 1e-14
    assert diag["gradient_projection_weighted_rms"] < 1e-9


def test_random_gradient_matches_oracle_and_dense_solve():
    pos, faces, area = uniform_graph(7, spacing=0.3)
    g = np.random.default_rng(3).standard_normal(pos.shape)
    p, diag = PP.recover_pressure(pos, faces, area, g, tolerance=1e-13)
    o, odiag = OP.recover_pressure(pos, faces, area, g, tolerance=1e-13)
    d = OP.recover_pressure_dense(pos, faces, area, g)
    scale = np.abs(d).max()
    print(f"[gpu] random gradient: vs oracle {np.abs(p - o).max() / scale:.2e}, vs dense "
          f"{np.abs(p - d).max() / scale:.2e}, iterations {diag['iterations']}", flush=True)
    assert np.abs(p - o).max() <= 1e-9 * scale and np.abs(p - d).max() <= 1e-9 * scale
    assert np.isclose(diag["gradient_projection_weighted_rms"], odiag["gradient_projection_weighted_rms"],
                      rtol=1e-6)


def test_run_snapshot_pressure_matches_oracle(tmp_path):
    cfg = golden_config("cpu_oscillating_sphere_smoke")
    cfg["runtime"]["device"] = "hip"
    cfg["solver"]["sparse"] = False
    cfg["output"]["save_field"] = True
    HC.run_case(HC.validate_config(cfg), tmp_path / "run", quiet=True)
    field = tmp_path / "run" / "final_field.npz"
    out = tmp_path / "p.npz"
    assert PP.main([str(field), "--out", str(out)]) == 0
    with np.load(out, allow_pickle=False) as r:
        p, diag = r["pressure"], json.loads(str(r["diagnostics"]))
    with np.load(field, allow_pickle=False) as s:
        o, _ = OP.recover_pressure(s["cpos"], s["face_cells"], s["face_area"], s["momentum_residual"] / float(s["dt"]))
    scale = np.abs(o).max()
    print(f"[gpu] snapshot pressure: {p.shape[0]} cells, vs oracle {np.abs(p - o).max() / scale:.2e}, {diag}",
          flush=True)
    assert np.abs(p - o).max() <= 1e-8 * scale
