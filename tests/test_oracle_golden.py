"""Oracle trajectories against the reference's own short runs (tests/golden/traj_*.npz).

The reference ran the bundled example cases with its CSR (``solver.sparse``)
CG on the CPU; the oracle runs the same configuration.  Iteration counts must
agree exactly and the state to round-off.
"""
import numpy as np
import pytest

from oracle.pyib_ref import case as OC
from tests.helpers import golden, golden_config, rel


@pytest.mark.parametrize("name,steps", [("cpu_fixed_sphere_smoke", 5), ("cpu_oscillating_sphere_smoke", 4),
                                        ("cpu_fixed_sphere_re100_d8_gate", 3)])
def test_oracle_reproduces_reference_trajectory(name, steps):
    cfg = golden_config(name)
    rt = OC.build_runtime(cfg)
    s, s_old, vc, rows = OC.run(rt, steps)
    with golden(f"traj_{name}.npz") as g:
        cols = [str(c) for c in g["columns"]]
        hist = g["history"]
        it = cols.index("helmholtz_iterations")
        assert [r["helmholtz_iterations"] for r in rows] == [int(v) for v in hist[:, it]]
        assert rel(s, g["s"]) < 1e-12
        assert rel(rt.driver.ib_force_L, g["ib_force_L"]) < 1e-10
        for k, col in enumerate(cols):
            if col in ("milliseconds_per_step",):
                continue
            ours = np.array([r[col] for r in rows])
            scale = max(np.abs(hist[:, k]).max(), 1e-300)
            assert np.abs(ours - hist[:, k]).max() <= 1e-9 * scale + 1e-15, col
        if "vc" in g.files:
            assert rel(vc[:rt.core.ncell], g["vc"]) < 1e-12
            assert rel(rt.driver.ib_force_E, g["ib_force_E"]) < 1e-10


@pytest.mark.parametrize("name,steps", [("cpu_fixed_sphere_smoke", 5), ("cpu_oscillating_sphere_smoke", 4),
                                        ("cpu_fixed_sphere_re100_d8_gate", 3)])
def test_matrix_free_oracle_meets_the_history_rules(name, steps):
    """The oracle's matrix-free path (the one the HIP path is compared with; the reference ran
    its CSR matrix) meets every per-column rule of tests.helpers.compare_history against the
    reference's history, and a planted error in one column is caught."""
    from tests.helpers import compare_history
    cfg = golden_config(name)
    rt = OC.build_runtime(cfg, sparse=False)
    rows = OC.run(rt, steps)[3]
    with golden(f"traj_{name}.npz") as g:
        cols = [str(c) for c in g["columns"]]
        hist = g["history"]
    report = compare_history(rows, hist, cols)
    print(name, {k: f"{v:.2g}" for k, v in report.items() if v > 1e-3})
    bad = [dict(r) for r in rows]
    bad[-1]["slip_max"] *= 1.0 + 1e-8
    with pytest.raises(AssertionError, match="slip_max"):
        compare_history(bad, hist, cols)
