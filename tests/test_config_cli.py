"""Format layer without a GPU: schema validation, fingerprints, CLI error contract.

Pins: the reference's own D32/D40 case files and the config fingerprints it
recorded for them (checks/pressure-validation-20260925.zip, kept as
tests/golden/reference_runs.json).
"""
import copy
import json
import os
import subprocess
import sys

import pytest

import pyib_hip  # noqa: F401
import pyib_hip.case as HC
from tests.helpers import GOLDEN, ROOT

REF = json.load(open(os.path.join(GOLDEN, "reference_runs.json")))


@pytest.mark.parametrize("label", ["D32", "D40"])
def test_config_fingerprint_matches_reference(label):
    cfg = HC.validate_config(copy.deepcopy(REF[label]["case"]))
    assert HC.config_fingerprint(cfg) == REF[label]["config_fingerprint"]


def test_additive_schema():
    base = copy.deepcopy(REF["D32"]["case"])
    base["runtime"] = {"device": "hip", "precision": "mixed"}
    base["mesh"] = {"kind": "nested_box", "cells_per_diameter": 32, "half_extents_d": [3, 2.5, 2.5], "levels": 3}
    cfg = HC.validate_config(base)
    assert cfg["mesh"]["levels"] == 3 and cfg["runtime"]["precision"] == "mixed"


@pytest.mark.parametrize("patch,msg", [
    ({"mesh": {"kind": "sphere_concentric", "cells_per_diameter": 12}}, "multiple of 8"),
    ({"runtime": {"device": "cpu", "precision": "mixed"}}, "mixed"),
    ({"solver": {"typo": 1}}, "unknown solver"),
    ({"ib": {"alpha": 0.5}}, "alpha=1.0"),
    ({"mesh": {"kind": "uniform_box", "half_extents_d": [1, 1, 1]}}, "nested_box"),
    ({"time": {"steps": 0}}, "positive integer"),
])
def test_rejections(patch, msg):
    raw = copy.deepcopy(REF["D32"]["case"])
    for k, v in patch.items():
        raw.setdefault(k, {}).update(v)
    with pytest.raises(HC.CaseConfigError, match=msg):
        HC.validate_config(raw)


def test_cli_reports_config_errors(tmp_path):
    raw = copy.deepcopy(REF["D32"]["case"])
    raw["ib"]["force_mode"] = "nope"
    case = tmp_path / "bad.json"
    case.write_text(json.dumps(raw))
    r = subprocess.run([sys.executable, "-m", "pyib_hip.cli", str(case), "--outdir", str(tmp_path / "out"),
                        "--device", "hip"], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 2
    assert "CFDAgent-IB error:" in r.stderr and "force_mode" in r.stderr


def test_cli_refuses_non_hip_device(tmp_path):
    case = tmp_path / "c.json"
    case.write_text(json.dumps(REF["D32"]["case"]))
    r = subprocess.run([sys.executable, "-m", "pyib_hip.cli", str(case), "--outdir", str(tmp_path / "o"),
                        "--device", "cpu"], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 2 and "runtime.device='hip'" in r.stderr


def test_reference_mesh_counts_closed_form():
    import pyib_hip.mesh as PM
    for label, ddx in (("D32", 32), ("D40", 40)):
        n = ddx
        shape = [(2 * n,) * 3] * 5
        c, f, e = PM.closed_form_counts(shape, 5)
        m = REF[label]["mesh"]
        assert (c, f, e) == (m["total_cells"], m["faces"], m["edges"])
