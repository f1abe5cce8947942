"""The artifact validator's host rules (no GPU).

The history rules accept the reference's own sampled histories (the short CPU runs in
tests/golden/traj_*.npz, written by the reference's runner) and reject each planted fault;
the mesh record of every mesh kind checks against the closed forms; the prescribed marker
state is the kinematics module's.  A full run directory is validated on the GPU
(tests/test_gpu_runner.py)."""
import copy
import csv

import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.case as HC
import pyib_hip.validate as V
from pyib_hip.kinematics import Oscillation
from tests.helpers import golden, golden_config


def _rows(name):
    with golden(f"traj_{name}.npz") as g:
        cols = [str(c) for c in g["columns"]]
        return [{c: (int(v) if c in V.INTEGER_COLUMNS else float(v)) for c, v in zip(cols, r)} for r in g["history"]]


def _check(rows, cfg):
    per = 1 if cfg["ib"]["force_mode"] == "accumulated_explicit" else int(cfg["time"]["inner_iterations"])
    V.check_history(rows, cfg, rows[-1]["step"], float(cfg["time"]["dt"]), per, 1e-9)


@pytest.mark.parametrize("name", ["cpu_fixed_sphere_smoke", "cpu_oscillating_sphere_smoke",
                                  "cpu_fixed_sphere_re100_d8_gate"])
def test_reference_histories_meet_the_rules(name):
    cfg = HC.validate_config(golden_config(name))
    cfg["output"]["sample_every"] = 1
    cfg["output"]["checkpoint_every"] = 0
    _check(_rows(name), cfg)


@pytest.mark.parametrize("fault,match", [
    ("step", "strictly increasing"), ("time", "step\\*dt"), ("slip", "negative"), ("balance", "force balance"),
    ("updates", "ib_force_updates"), ("gap", "omits required sampled steps"), ("cg", "unconverged CG"),
    ("load", "interior-momentum correction"), ("start", "begin at step 1"),
])
def test_planted_history_faults_are_rejected(fault, match):
    name = "cpu_fixed_sphere_smoke"
    cfg = HC.validate_config(golden_config(name))
    cfg["output"].update(sample_every=1, checkpoint_every=0)
    cfg["solver"]["fail_on_cg_nonconvergence"] = True
    rows = copy.deepcopy(_rows(name))
    if fault == "step":
        rows[2]["step"] = rows[1]["step"]
    elif fault == "time":
        rows[2]["time"] *= 1.0 + 1e-9
    elif fault == "slip":
        rows[1]["slip_max"] = -1.0
    elif fault == "balance":
        rows[3]["force_balance_relerr"] = 2e-9
    elif fault == "updates":
        rows[0]["ib_force_updates"] = 2
    elif fault == "gap":
        del rows[2]
    elif fault == "cg":
        rows[1]["helmholtz_tolerance_ratio"] = 1.01
    elif fault == "load":
        rows[4]["body_force_x"] += 1e-6
    elif fault == "start":
        rows = rows[1:]
    with pytest.raises(V.ValidationError, match=match):
        _check(rows, cfg)


def test_history_reader_round_trip(tmp_path):
    rows = _rows("cpu_oscillating_sphere_smoke")
    p = tmp_path / "history.csv"
    with p.open("w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=HC.HISTORY_COLUMNS)
        w.writeheader()
        w.writerows(rows)
    back = V.read_history(p)
    assert [r["step"] for r in back] == [r["step"] for r in rows]
    assert all(np.isclose(a["Cd"], b["Cd"], rtol=1e-15) for a, b in zip(back, rows))
    text = p.read_text().splitlines()
    p.write_text("\n".join([text[0].replace("Cd", "CD")] + text[1:]) + "\n")
    with pytest.raises(V.ValidationError, match="header"):
        V.read_history(p)
    p.write_text("\n".join(text[:2] + [text[2].replace(text[2].split(",")[3], "nan", 1)]) + "\n")
    with pytest.raises(V.ValidationError, match="non-finite|not numeric"):
        V.read_history(p)


@pytest.mark.parametrize("mesh", [
    {"kind": "uniform_box", "dx": 0.25, "bounds": [[-1, 1], [-1, 1], [-1, 1]]},
    {"kind": "nested_box", "cells_per_diameter": 8, "half_extents_d": [1.5, 1.25, 1.25], "levels": 2},
    {"kind": "sphere_wake", "cells_per_diameter": 8},
])
def test_mesh_record_checks_against_closed_forms(mesh):
    raw = {"name": "m", "case": {"kind": "fixed_sphere"}, "runtime": {"device": "hip"}, "mesh": mesh,
           "physics": {"reynolds": 100.0, "inflow_speed": 1.0}, "ib": {"marker_spacing": 0.0625}}
    cfg = HC.validate_config(raw)
    m, dx, summary = HC.build_mesh(cfg)
    summary.update(faces=int(m.faces_to_cells.shape[0]), edges=int(m.edges_to_nodes.shape[0]),
                   markers=V.marker_count(cfg))
    cells, faces, edges, markers, widths = V._mesh_counts(cfg, summary)
    assert (cells, faces, edges) == (m.cells.shape[0], m.faces_to_cells.shape[0], m.edges_to_nodes.shape[0])
    assert widths.shape == (cells,) and widths[0] == dx
    bad = dict(summary, edges=summary["edges"] + 1)
    with pytest.raises(V.ValidationError, match="closed form"):
        V._mesh_counts(cfg, bad)
    bad = dict(summary, markers=summary["markers"] + 1)
    with pytest.raises(V.ValidationError, match="markers"):
        V._mesh_counts(cfg, bad)


def test_prescribed_markers_follow_the_kinematics():
    raw = {"name": "o", "case": {"kind": "oscillating_sphere"}, "runtime": {"device": "hip"},
           "mesh": {"kind": "uniform_box", "dx": 0.25, "bounds": [[-1, 1], [-1, 1], [-1, 1]]},
           "physics": {"reynolds": 0.2, "inflow_speed": 0.0},
           "ib": {"marker_spacing": 0.25, "motion": {"kind": "oscillation", "axis": [1, 0, 0], "amplitude": 0.05,
                                                      "frequency": 0.5, "phase": 0.3}}}
    cfg = HC.validate_config(raw)
    n = V.marker_count(cfg)
    ref = HC.fibonacci_sphere(n, 0.25, [0, 0, 0])
    osc = Oscillation(ref, axis=[1, 0, 0], A=0.05, freq=0.5, phase=0.3)
    for t in (0.0, 0.37, 1.9):
        m, v = V.prescribed_markers(cfg, n, t)
        m2, v2 = osc(t)
        np.testing.assert_allclose(m, m2, rtol=0, atol=1e-15)
        np.testing.assert_allclose(v, v2, rtol=0, atol=1e-15)


def test_missing_directory_and_artifacts(tmp_path):
    with pytest.raises(V.ValidationError, match="does not exist"):
        V.validate_result(tmp_path / "absent")
    with pytest.raises(V.ValidationError, match="missing required result artifacts"):
        V.validate_result(tmp_path)
    assert V.main([str(tmp_path)]) == 1
