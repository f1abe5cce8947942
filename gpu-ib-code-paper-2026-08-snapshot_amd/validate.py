"""Acceptance of the artifacts a HIP run writes (``run_case`` / ``pyib-run --device hip``).

Counterpart of the reference's ``validate_ib_result.py`` (``validate_result`` 1269-1956,
history rules ``_validate_history`` 1168-1242, CG evidence 1111-1118).  The reference's
validator refuses any device but cpu/cuda (416) and any source identity but its own files
(1398-1404), so a HIP run needs its own: this one accepts ``runtime.device = "hip"`` with
``precision`` float64 or mixed, and the source identity of this package
(``case.source_hashes``).  The acceptance contract is the reference's: complete artifacts,
one run identity across them, fingerprints that recompute, the prescribed body state at the
checkpoint time, the closed-body load algebra ``body = -fluid + d/dt(interior)``, the sampled
history (steps, times, update counts, force balance, strict CG evidence), and, when the final
field is saved, the loads recomputed from the fields.  Mixed precision keeps every field and
load in float64, so both precisions are held to the float64 tolerances.

    python -m pyib_hip.validate RESULT_DIR [--force-balance-atol X]

prints ``VALID: ...`` and returns 0, or ``INVALID: reason`` on stderr and returns 1 (2 if
the validator itself fails), as the reference's command does.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import sys
import uuid
from datetime import datetime
from pathlib import Path

import numpy as np

from . import case as HC
from . import mesh as M

FIELD_DTYPE = np.dtype(np.float64)       # float64 and mixed runs both persist float64
FORCE_BALANCE_ATOL = 1.0e-9
LOAD_TOL = 1.0e-12
DIAG_RTOL, DIAG_ATOL = 1.0e-10, 1.0e-12
LOAD_VECTORS = ("force_fluid", "moment_fluid", "body_force", "body_moment", "interior_momentum",
                "interior_angular_momentum", "interior_momentum_rate", "interior_angular_momentum_rate")
INTEGER_COLUMNS = {"step", "ib_force_updates", "helmholtz_iterations"}
NONNEGATIVE_COLUMNS = {"helmholtz_relative_residual", "helmholtz_tolerance_ratio", "slip_max", "slip_rms",
                       "speed_max", "cfl", "mass_residual_inf", "force_balance_relerr", "ib_force_updates",
                       "helmholtz_iterations", "milliseconds_per_step", "interior_indicator_volume"}
METADATA_KEYS = {"result_schema", "run_id", "created_utc", "case_name", "case_kind", "backend", "precision",
                 "python", "array_backend_version", "config_fingerprint", "layout_fingerprint", "config", "mesh",
                 "dt", "target_steps", "start_step", "resumed_from", "output_directory", "source_sha256",
                 "closed_body_load", "resume_count", "setup_seconds", "completed_utc", "completed_step", "complete"}
CHECKPOINT_KEYS = {"checkpoint_schema", "update_phase", "run_id", "config_fingerprint", "layout_fingerprint", "step",
                   "physical_time", "dt", "previous_dt", "precision", "device", "source_hashes_json", "s", "s_old",
                   "markers", "marker_velocity", "ib_force_mode", "ib_force_E", "ib_force_L",
                   "ib_force_update_count", "interior_indicator_volume", *LOAD_VECTORS}
FINAL_KEYS = (CHECKPOINT_KEYS - {"checkpoint_schema", "update_phase", "physical_time", "previous_dt"}) | {
    "result_schema", "time", "vc", "cpos", "cell_volume", "marker_volume_weight", "momentum_residual",
    "helmholtz_relative_residual", "helmholtz_tolerance_ratio", "face_cells", "face_area", "index_cell",
    "lattice_origin", "lattice_spacing", "sphere_center", "sphere_radius", "density", "kinematic_viscosity"}
CLOSED_BODY_LOAD = {
    "method": "Uhlmann interior-fluid momentum correction",
    "force": "-force_fluid + backward_difference(interior_momentum)",
    "moment": "-moment_fluid + backward_difference(interior_angular_momentum)",
    "indicator": "analytic sphere with one-cell linear smoothing",
}


class ValidationError(RuntimeError):
    """A persisted HIP result violates the acceptance contract."""


def require(ok, message):
    if not ok:
        raise ValidationError(message)


# ------------------------------------------------------------------ primitive readers
def _strict_json(text, label):
    def no_constant(token):
        raise ValidationError(f"{label} contains the non-finite JSON constant {token}")

    def unique(pairs):
        out = {}
        for k, v in pairs:
            require(k not in out, f"{label} repeats the key {k!r}")
            out[k] = v
        return out
    try:
        return json.loads(text, parse_constant=no_constant, object_pairs_hook=unique)
    except json.JSONDecodeError as exc:
        raise ValidationError(f"{label} is not valid JSON: {exc}") from exc


def _utc(value, label):
    require(isinstance(value, str), f"{label} must be an ISO-8601 string")
    try:
        t = datetime.fromisoformat(value)
    except ValueError as exc:
        raise ValidationError(f"{label} is not ISO-8601") from exc
    require(t.tzinfo is not None and t.utcoffset().total_seconds() == 0, f"{label} must be UTC")
    return t


def _sha(value, label):
    require(isinstance(value, str) and len(value) == 64 and all(c in "0123456789abcdef" for c in value),
            f"{label} must be a lower-case sha256 digest")
    return value


def _int(value, label, minimum=None):
    require(isinstance(value, int) and not isinstance(value, bool), f"{label} must be an integer")
    require(minimum is None or value >= minimum, f"{label} must be >= {minimum}")
    return int(value)


def _npz(path, keys):
    try:
        with np.load(path, allow_pickle=False) as a:
            got = set(a.files)
            require(got == keys, f"{path.name} fields differ: missing {sorted(keys - got)}, "
                                 f"unexpected {sorted(got - keys)}")
            return {k: a[k] for k in a.files}
    except ValidationError:
        raise
    except (OSError, ValueError, EOFError) as exc:
        raise ValidationError(f"cannot read {path.name}: {exc}") from exc


def _scalar(a, label, kind):
    a = np.asarray(a)
    require(a.shape == (), f"{label} must be a scalar")
    v = a.item()
    if kind is str:
        require(a.dtype.kind == "U", f"{label} must be a string")
        return str(v)
    if kind is int:
        require(a.dtype.kind in "iu", f"{label} must be an integer")
        return int(v)
    require(a.dtype.kind == "f" and math.isfinite(float(v)), f"{label} must be a finite float")
    return float(v)


def _field(a, label, shape):
    require(a.dtype == FIELD_DTYPE, f"{label} dtype {a.dtype} != {FIELD_DTYPE}")
    require(a.shape == tuple(shape), f"{label} shape {a.shape} != {tuple(shape)}")
    require(bool(np.all(np.isfinite(a))), f"{label} contains non-finite values")


def _time_tol(step, dt):
    return max(1.0e-12, 16.0 * np.finfo(np.float64).eps * max(abs(step * dt), dt) * max(step, 1))


# ------------------------------------------------------------------ physics of the case
def reference_speed(config):
    u = abs(float(config["physics"]["inflow_speed"]))
    if u == 0.0:
        mo = config["ib"]["motion"]
        u = 2.0 * math.pi * float(mo["frequency"]) * float(mo["amplitude"])
    return u


def marker_count(config):
    r, h = float(config["ib"]["radius"]), float(config["ib"]["marker_spacing"])
    return max(12, int(round(4.0 * math.pi * r * r / (h * h))))


def prescribed_markers(config, n, t):
    """Fibonacci markers at the prescribed body state (ib_case.py:554-563, 723-763)."""
    ib = config["ib"]
    centre, radius = np.asarray(ib["center"], float), float(ib["radius"])
    ref = HC.fibonacci_sphere(n, radius, centre)
    mo = ib["motion"]
    if mo["kind"] == "none":
        return ref, np.zeros_like(ref)
    axis = np.asarray(mo["axis"], float)
    if mo["kind"] == "rotation":
        return ref, HC.rotating_sphere_velocity(ref, centre, axis, float(mo["spin_ratio"]), radius,
                                                reference_speed(config))
    w = 2.0 * math.pi * float(mo["frequency"])
    ang = w * t + float(mo["phase"])
    a = float(mo["amplitude"])
    return ref + a * math.sin(ang) * axis, np.tile(a * w * math.cos(ang) * axis, (n, 1))


def _load_algebra(loads, label):
    require(np.allclose(loads["body_force"], -loads["force_fluid"] + loads["interior_momentum_rate"],
                        rtol=LOAD_TOL, atol=LOAD_TOL), f"{label} body force violates the interior-momentum correction")
    require(np.allclose(loads["body_moment"], -loads["moment_fluid"] + loads["interior_angular_momentum_rate"],
                        rtol=LOAD_TOL, atol=LOAD_TOL),
            f"{label} body moment violates the interior-angular-momentum correction")


def _mesh_counts(config, mesh_md):
    """(cells, faces, edges, markers) from the metadata, each checked against the closed forms of
    the case's boxes; also the per-layer cell widths."""
    kind = config["mesh"]["kind"]
    require(mesh_md.get("kind") == kind, "metadata.mesh.kind disagrees with config")
    if kind == "uniform_box":
        bounds = np.asarray(config["mesh"]["bounds"], float)
        dx = float(config["mesh"]["dx"])
        require(mesh_md.get("dx") == dx, "metadata.mesh.dx disagrees with config")
        require(np.array_equal(np.asarray(mesh_md.get("bounds"), float), bounds),
                "metadata.mesh.bounds disagrees with config")
        shapes = [tuple(int(v) for v in np.rint((bounds[:, 1] - bounds[:, 0]) / dx))]
        cells = _int(mesh_md.get("cells"), "metadata.mesh.cells", 1)
        widths = [dx]
    else:
        ddx = _int(mesh_md.get("D_over_dx"), "metadata.mesh.D_over_dx", 1)
        require(ddx == int(config["mesh"]["cells_per_diameter"]), "metadata.mesh.D_over_dx disagrees with config")
        require(mesh_md.get("diameter") == 2.0 * float(config["ib"]["radius"]),
                "metadata.mesh.diameter disagrees with the IB radius")
        require(mesh_md.get("dx_fine") == M.DIAMETER / ddx, "metadata.mesh.dx_fine disagrees with diameter/D_over_dx")
        shapes = [tuple(int(v) for v in s) for s in mesh_md.get("layer_shape", [])]
        require(bool(shapes), "metadata.mesh.layer_shape is missing")
        for k in range(1, len(shapes)):
            require(all(shapes[k - 1][a] % 2 == 0 for a in range(3)), "a finer layer has an odd cell count")
        cells = _int(mesh_md.get("total_cells"), "metadata.mesh.total_cells", 1)
        layer = [int(v) for v in mesh_md.get("layer_cells", [])]
        want = [int(np.prod(shapes[0]))] + [int(np.prod(shapes[k])) - int(np.prod([v // 2 for v in shapes[k - 1]]))
                                            for k in range(1, len(shapes))]
        require(layer == want, "metadata.mesh.layer_cells disagrees with the layer shapes")
        require(sum(layer) == cells, "metadata.mesh.layer_cells do not sum to total_cells")
        widths = [M.DIAMETER / ddx * 2 ** k for k in range(len(shapes))]
    c, f, e = M.closed_form_counts(shapes, len(shapes))
    require(cells == c, "metadata.mesh cell count disagrees with the closed form of its boxes")
    faces = _int(mesh_md.get("faces"), "metadata.mesh.faces", 1)
    edges = _int(mesh_md.get("edges"), "metadata.mesh.edges", 1)
    require(faces == f and edges == e, "metadata.mesh face/edge counts disagree with the closed form of its boxes")
    markers = _int(mesh_md.get("markers"), "metadata.mesh.markers", 1)
    require(markers == marker_count(config), "metadata.mesh.markers disagrees with the IB geometry")
    layer_cells = [int(v) for v in mesh_md.get("layer_cells", [cells])]
    return cells, faces, edges, markers, np.repeat(np.asarray(widths, float), layer_cells)


# ------------------------------------------------------------------ history
def read_history(path):
    rows = []
    try:
        with Path(path).open("r", newline="", encoding="utf-8") as fh:
            rd = csv.reader(fh)
            header = next(rd, None)
            require(header is not None, "history.csv is empty")
            require(list(header) == list(HC.HISTORY_COLUMNS), "history.csv header/schema mismatch")
            for line, vals in enumerate(rd, start=2):
                require(len(vals) == len(header), f"history.csv line {line} has {len(vals)} fields")
                row = {}
                for name, raw in zip(header, vals):
                    raw = raw.strip()
                    require(bool(raw), f"history.csv line {line} has an empty {name}")
                    try:
                        row[name] = int(raw) if name in INTEGER_COLUMNS else float(raw)
                    except ValueError as exc:
                        raise ValidationError(f"history.csv line {line} {name} is not numeric") from exc
                    require(name in INTEGER_COLUMNS or math.isfinite(row[name]),
                            f"history.csv line {line} {name} is non-finite")
                rows.append(row)
    except ValidationError:
        raise
    except (OSError, UnicodeError, csv.Error) as exc:
        raise ValidationError(f"cannot read history.csv: {exc}") from exc
    require(bool(rows), "history.csv contains no data rows")
    return rows


def _vec(row, stem):
    return np.asarray([row[f"{stem}_{a}"] for a in "xyz"], float)


def check_history(rows, config, completed, dt, updates_per_step, balance_atol):
    """ib_case.py run-time invariants as persisted, validate_ib_result.py:1168-1242 and 1111-1118."""
    require(rows[0]["step"] == 1, "history.csv must begin at step 1")
    prev_step, prev_t = 0, -math.inf
    for line, r in enumerate(rows, start=2):
        step, t = int(r["step"]), float(r["time"])
        require(step > prev_step, f"history.csv step is not strictly increasing at line {line}")
        require(t > prev_t, f"history.csv time is not strictly increasing at line {line}")
        require(step <= completed, f"history.csv line {line} exceeds completed_step")
        require(abs(t - step * dt) <= _time_tol(step, dt), f"history.csv line {line} time disagrees with step*dt")
        for c in NONNEGATIVE_COLUMNS:
            require(r[c] >= 0.0, f"history.csv line {line} {c} is negative")
        require(r["interior_indicator_volume"] > 0.0, f"history.csv line {line} interior_indicator_volume <= 0")
        _load_algebra({k: _vec(r, k) for k in LOAD_VECTORS}, f"history.csv line {line}")
        require(r["ib_force_updates"] == updates_per_step,
                f"history.csv line {line} ib_force_updates != {updates_per_step}")
        require(r["force_balance_relerr"] <= balance_atol,
                f"history.csv line {line} force balance {r['force_balance_relerr']:.6e} exceeds {balance_atol:.6e}")
        if config["solver"]["fail_on_cg_nonconvergence"]:
            require(r["helmholtz_tolerance_ratio"] <= 1.0, "strict run contains unconverged CG history")
        prev_step, prev_t = step, t
    require(rows[-1]["step"] == completed, "history.csv final step disagrees with metadata")
    out = config["output"]
    need = set(range(1, min(completed, 3) + 1)) | set(range(int(out["sample_every"]), completed + 1,
                                                            int(out["sample_every"]))) | {completed}
    if int(out["checkpoint_every"]):
        need |= set(range(int(out["checkpoint_every"]), completed + 1, int(out["checkpoint_every"])))
    missing = sorted(need - {int(r["step"]) for r in rows})
    require(not missing, f"history.csv omits required sampled steps: {missing}")


# ------------------------------------------------------------------ the whole result
def validate_result(result_directory, *, force_balance_atol=None):
    """Validate a complete HIP result directory; returns a summary dict or raises ValidationError."""
    root = Path(result_directory).expanduser().resolve()
    require(root.is_dir(), f"result directory does not exist: {root}")
    p = {k: root / n for k, n in (("md", "metadata.json"), ("hist", "history.csv"), ("ck", "checkpoint.npz"),
                                  ("final", "final_field.npz"))}
    missing = [p[k].name for k in ("md", "hist", "ck") if not p[k].is_file()]
    require(not missing, f"missing required result artifacts: {missing}")
    balance_atol = FORCE_BALANCE_ATOL if force_balance_atol is None else float(force_balance_atol)
    require(math.isfinite(balance_atol) and balance_atol >= 0.0, "force_balance_atol must be finite and >= 0")

    # ---- metadata and the configuration it carries
    md = _strict_json(p["md"].read_text(encoding="utf-8"), "metadata.json")
    require(isinstance(md, dict), "metadata.json must hold an object")
    extra, absent = set(md) - METADATA_KEYS - {"last_resume_utc"}, METADATA_KEYS - set(md)
    require(not extra and not absent, f"metadata keys differ: missing {sorted(absent)}, unexpected {sorted(extra)}")
    require(md["result_schema"] == HC.RESULT_SCHEMA_VERSION, f"unsupported result_schema; expected "
                                                             f"{HC.RESULT_SCHEMA_VERSION}")
    created, completed_t = _utc(md["created_utc"], "metadata.created_utc"), _utc(md["completed_utc"],
                                                                               "metadata.completed_utc")
    require(completed_t >= created, "metadata.completed_utc predates created_utc")
    try:
        run_id = str(uuid.UUID(str(md["run_id"])))
    except ValueError as exc:
        raise ValidationError("metadata.run_id is not a UUID") from exc
    require(run_id == md["run_id"], "metadata.run_id is not in canonical form")
    layout_fp = _sha(md["layout_fingerprint"], "metadata.layout_fingerprint")
    try:
        config = HC.validate_config(md["config"])
    except HC.CaseConfigError as exc:
        raise ValidationError(f"metadata.config is not a valid case: {exc}") from exc
    require(config == md["config"], "metadata.config is not in normalised form")
    require(config["runtime"]["device"] == "hip", "config.runtime.device must be hip")
    precision = config["runtime"]["precision"]
    require(precision in ("float64", "mixed"), "config.runtime.precision must be float64 or mixed")
    require(md["precision"] == precision, "metadata precision disagrees with config")
    require(md["backend"] == HC.BACKEND, "metadata backend disagrees with the HIP path")
    require(md["case_name"] == config["name"] and md["case_kind"] == config["case"]["kind"],
            "metadata case name/kind disagrees with config")
    cbl = md["closed_body_load"]
    require(isinstance(cbl, dict) and set(cbl) == set(CLOSED_BODY_LOAD) | {"moment_origin"},
            "metadata.closed_body_load schema mismatch")
    require(all(cbl[k] == v for k, v in CLOSED_BODY_LOAD.items()), "metadata closed-body load definition "
                                                                     "is unsupported")
    require(np.array_equal(np.asarray(cbl["moment_origin"], float), np.zeros(3)),
            "metadata closed-body moment origin must be the global origin")
    require(md["complete"] is True, "strict acceptance requires a complete run")
    dt = float(config["time"]["dt"])
    require(md["dt"] == dt, "metadata.dt disagrees with config.time.dt")
    target = _int(md["target_steps"], "metadata.target_steps", 1)
    start = _int(md["start_step"], "metadata.start_step", 0)
    completed = _int(md["completed_step"], "metadata.completed_step", 1)
    resumes = _int(md["resume_count"], "metadata.resume_count", 0)
    require(target == config["time"]["steps"] and completed == target,
            "metadata target/completed steps disagree with config")
    require(start < completed, "metadata.start_step must be before completed_step")
    if start == 0:
        require(md["resumed_from"] is None and resumes == 0 and "last_resume_utc" not in md,
                "a fresh run's metadata must not carry resume information")
    else:
        require(isinstance(md["resumed_from"], str) and Path(md["resumed_from"]).name == "checkpoint.npz",
                "metadata.resumed_from must name the run's checkpoint.npz")
        require(resumes >= 1 and "last_resume_utc" in md, "resumed metadata lacks its resume record")
        require(created <= _utc(md["last_resume_utc"], "metadata.last_resume_utc") <= completed_t,
                "metadata.last_resume_utc lies outside the run lifetime")
    require(isinstance(md["setup_seconds"], (int, float)) and md["setup_seconds"] >= 0.0,
            "metadata.setup_seconds must be a non-negative number")
    fp = _sha(md["config_fingerprint"], "metadata.config_fingerprint")
    require(fp == HC.config_fingerprint(config), "metadata config fingerprint is incorrect")
    cells, faces, edges, markers, widths = _mesh_counts(config, md["mesh"])

    # ---- source identity: this package's numerical sources as they are now
    src = md["source_sha256"]
    require(isinstance(src, dict) and set(src) == set(HC.source_hashes()), "metadata.source_sha256 key set mismatch")
    for name, digest in HC.source_hashes().items():
        require(_sha(src[name], f"metadata.source_sha256.{name}") == digest, f"numerical source hash mismatch "
                                                                              f"for {name}")

    # ---- checkpoint
    save_field = bool(config["output"]["save_field"])
    require(p["final"].is_file() == save_field, "final_field.npz must exist exactly when output.save_field is true")
    ck = _npz(p["ck"], CHECKPOINT_KEYS)
    require(_scalar(ck["checkpoint_schema"], "checkpoint.checkpoint_schema", int) == HC.CHECKPOINT_SCHEMA_VERSION,
            "unsupported checkpoint schema")
    require(_scalar(ck["update_phase"], "checkpoint.update_phase", str) == "post_step",
            "checkpoint update_phase must be post_step")
    require(_scalar(ck["run_id"], "checkpoint.run_id", str) == run_id, "checkpoint run_id disagrees with metadata")
    require(_scalar(ck["layout_fingerprint"], "checkpoint.layout_fingerprint", str) == layout_fp,
            "checkpoint layout fingerprint disagrees with metadata")
    require(_scalar(ck["config_fingerprint"], "checkpoint.config_fingerprint", str) == fp,
            "checkpoint config fingerprint mismatch")
    step = _scalar(ck["step"], "checkpoint.step", int)
    require(step == completed, "checkpoint step disagrees with metadata")
    t_ck = _scalar(ck["physical_time"], "checkpoint.physical_time", float)
    require(abs(t_ck - step * dt) <= _time_tol(step, dt), "checkpoint physical_time disagrees with step*dt")
    require(_scalar(ck["dt"], "checkpoint.dt", float) == dt and
            _scalar(ck["previous_dt"], "checkpoint.previous_dt", float) == dt, "checkpoint dt mismatch")
    require(_scalar(ck["precision"], "checkpoint.precision", str) == precision, "checkpoint precision mismatch")
    require(_scalar(ck["device"], "checkpoint.device", str) == "hip", "checkpoint device mismatch")
    mode = config["ib"]["force_mode"]
    require(_scalar(ck["ib_force_mode"], "checkpoint.ib_force_mode", str) == mode,
            "checkpoint IB force mode disagrees with config")
    ck_src = _strict_json(_scalar(ck["source_hashes_json"], "checkpoint.source_hashes_json", str),
                          "checkpoint.source_hashes_json")
    require(ck_src == src, "checkpoint source identity disagrees with metadata")
    per_step = 1 if mode == "accumulated_explicit" else int(config["time"]["inner_iterations"])
    upd = _scalar(ck["ib_force_update_count"], "checkpoint.ib_force_update_count", int)
    require(upd == completed * per_step, "checkpoint IB force update count disagrees with completed steps")
    for name, shape in (("s", (edges,)), ("s_old", (edges,)), ("markers", (markers, 3)),
                        ("marker_velocity", (markers, 3)), ("ib_force_E", (cells, 3)), ("ib_force_L", (markers, 3))):
        _field(ck[name], f"checkpoint.{name}", shape)
    loads = {}
    for name in LOAD_VECTORS:
        _field(ck[name], f"checkpoint.{name}", (3,))
        loads[name] = ck[name]
    loads["interior_indicator_volume"] = _scalar(ck["interior_indicator_volume"], "checkpoint.interior_indicator_volume",
                                                 float)
    require(loads["interior_indicator_volume"] > 0.0, "checkpoint interior indicator volume must be positive")
    _load_algebra(loads, "checkpoint")
    want_m, want_v = prescribed_markers(config, markers, t_ck)
    require(np.allclose(ck["markers"], want_m, rtol=0.0, atol=1e-12) and
            np.allclose(ck["marker_velocity"], want_v, rtol=0.0, atol=1e-12),
            "checkpoint markers disagree with the prescribed body state")

    # ---- history
    hist = read_history(p["hist"])
    check_history(hist, config, completed, dt, per_step, balance_atol)
    last = hist[-1]
    for name in LOAD_VECTORS:
        require(np.allclose(_vec(last, name), loads[name], rtol=LOAD_TOL, atol=LOAD_TOL),
                f"history final {name} disagrees with the checkpoint's loads")
    require(math.isclose(last["interior_indicator_volume"], loads["interior_indicator_volume"], rel_tol=LOAD_TOL,
                         abs_tol=LOAD_TOL), "history final interior indicator volume disagrees with the checkpoint")
    require(abs(last["time"] - t_ck) <= _time_tol(completed, dt), "history final time disagrees with checkpoint")
    summary = {"result_directory": str(root), "case_name": config["name"], "case_kind": config["case"]["kind"],
               "backend": md["backend"], "precision": precision, "completed_step": completed,
               "history_rows": len(hist), "force_balance_max": max(r["force_balance_relerr"] for r in hist),
               "force_balance_recomputed": None, "force_balance_atol": balance_atol, "config_fingerprint": fp,
               "coverage": "checkpoint_only"}
    if not save_field:
        return summary

    # ---- final field: the checkpoint's state plus what the loads are recomputed from
    fin = _npz(p["final"], FINAL_KEYS)
    require(_scalar(fin["result_schema"], "final_field.result_schema", int) == HC.RESULT_SCHEMA_VERSION,
            "unsupported final_field result_schema")
    for name in ("run_id", "layout_fingerprint", "config_fingerprint", "device", "precision", "ib_force_mode",
                 "source_hashes_json"):
        require(_scalar(fin[name], f"final_field.{name}", str) == _scalar(ck[name], name, str),
                f"final_field {name} disagrees with the checkpoint")
    require(_scalar(fin["step"], "final_field.step", int) == completed, "final_field step disagrees with metadata")
    t_fin = _scalar(fin["time"], "final_field.time", float)
    require(t_fin == t_ck, "final_field time disagrees with checkpoint")
    require(_scalar(fin["dt"], "final_field.dt", float) == dt, "final_field dt mismatch")
    require(_scalar(fin["ib_force_update_count"], "final_field.ib_force_update_count", int) == upd,
            "final_field IB force update count disagrees with checkpoint")
    for name in ("s", "s_old", "markers", "marker_velocity", "ib_force_E", "ib_force_L", "interior_indicator_volume",
                 *LOAD_VECTORS):
        require(fin[name].dtype == ck[name].dtype and np.array_equal(fin[name], ck[name]),
                f"{name} differs between checkpoint and final field")
    for name, shape in (("vc", (cells, 3)), ("momentum_residual", (cells, 3)), ("cpos", (cells, 3)),
                        ("cell_volume", (cells,)), ("marker_volume_weight", (markers,)), ("face_area", (faces,)),
                        ("lattice_origin", (3,)), ("sphere_center", (3,))):
        _field(fin[name], f"final_field.{name}", shape)
    fc = fin["face_cells"]
    require(fc.dtype.kind in "iu" and fc.shape == (faces, 2) and bool(np.all((fc[:, 0] >= 0) & (fc[:, 0] < cells)))
            and bool(np.all((fc[:, 1] >= -1) & (fc[:, 1] < cells))), "final_field.face_cells is not face connectivity")
    ic = fin["index_cell"]
    require(ic.dtype.kind in "iu" and ic.ndim == 3 and ic.size > 0 and bool(np.all((ic >= -1) & (ic < cells))),
            "final_field.index_cell is not a lattice of cell ids")
    require(bool(np.all(fin["face_area"] > 0.0)), "final_field.face_area must be positive")
    require(np.allclose(fin["sphere_center"], HC.sphere_center(config, t_fin), rtol=0.0, atol=1e-12),
            "final_field.sphere_center disagrees with the prescribed motion")
    rho, radius = float(config["physics"]["density"]), float(config["ib"]["radius"])
    nu = reference_speed(config) * 2.0 * radius / float(config["physics"]["reynolds"])
    dx_fine = float(widths[0]) if config["mesh"]["kind"] != "uniform_box" else float(config["mesh"]["dx"])
    for name, want in (("lattice_spacing", dx_fine), ("sphere_radius", radius), ("density", rho),
                       ("kinematic_viscosity", nu)):
        v = _scalar(fin[name], f"final_field.{name}", float)
        require(v > 0.0 and math.isclose(v, want, rel_tol=1e-12, abs_tol=0.0),
                f"final_field.{name} disagrees with the configuration")
    for name in ("helmholtz_relative_residual", "helmholtz_tolerance_ratio"):
        v = _scalar(fin[name], f"final_field.{name}", float)
        require(v >= 0.0 and v == last[name], f"final_field.{name} disagrees with history")
    # cell volumes: the width of each cell's layer, cubed (cells are numbered layer by layer)
    require(np.allclose(fin["cell_volume"], widths ** 3, rtol=1e-12, atol=0.0),
            "final_field.cell_volume disagrees with the mesh's layer spacings")
    ds = (4.0 * math.pi * radius ** 2 * dx_fine / markers) ** (1.0 / 3.0)
    require(np.allclose(fin["marker_volume_weight"], np.full(markers, ds ** 3), rtol=1e-12, atol=0.0),
            "final_field.marker_volume_weight disagrees with the IB quadrature")
    vol = fin["cell_volume"]
    fluid = (fin["ib_force_E"] * vol[:, None]).sum(axis=0)
    lag = (fin["ib_force_L"] * fin["marker_volume_weight"][:, None]).sum(axis=0)
    den = max(float(np.linalg.norm(lag)), float(np.linalg.norm(fluid)), 1e-300)
    balance = float(np.linalg.norm(fluid - lag)) / den
    require(np.allclose(fin["force_fluid"], fluid, rtol=LOAD_TOL, atol=LOAD_TOL),
            "persisted force_fluid disagrees with the Eulerian IB force field")
    moment = (np.cross(fin["cpos"], fin["ib_force_E"]) * vol[:, None]).sum(axis=0)
    require(np.allclose(fin["moment_fluid"], moment, rtol=LOAD_TOL, atol=LOAD_TOL),
            "persisted moment_fluid disagrees with the Eulerian IB force field")
    centre = np.asarray(HC.sphere_center(config, t_fin), float)
    dist = np.linalg.norm(fin["cpos"] - centre, axis=1)
    w = np.clip((radius + 0.5 * dx_fine - dist) / dx_fine, 0.0, 1.0) * vol
    lin = rho * (fin["vc"] * w[:, None]).sum(axis=0)
    ang = rho * (np.cross(fin["cpos"], fin["vc"]) * w[:, None]).sum(axis=0)
    require(np.allclose(fin["interior_momentum"], lin, rtol=LOAD_TOL, atol=LOAD_TOL) and
            np.allclose(fin["interior_angular_momentum"], ang, rtol=LOAD_TOL, atol=LOAD_TOL) and
            math.isclose(float(fin["interior_indicator_volume"]), float(w.sum()), rel_tol=LOAD_TOL, abs_tol=LOAD_TOL),
            "persisted interior momenta disagree with the final velocity field")
    require(balance <= balance_atol, f"recomputed final IB force balance {balance:.6e} exceeds {balance_atol:.6e}")
    require(math.isclose(last["force_balance_relerr"], balance, rel_tol=5e-3, abs_tol=1e-12),
            "history final force_balance_relerr disagrees with the persisted IB forces")
    fref = 0.5 * rho * reference_speed(config) ** 2 * math.pi * radius ** 2
    require(np.allclose(_vec(last, "body_force"), fin["body_force"], rtol=DIAG_RTOL, atol=DIAG_ATOL) and
            np.allclose([last["Cd"], last["Cy"], last["Cz"]], fin["body_force"] / fref, rtol=DIAG_RTOL,
                        atol=DIAG_ATOL), "history final force coefficients disagree with the persisted body load")
    summary.update(force_balance_recomputed=balance, coverage="full_field")
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("result_directory", type=Path)
    ap.add_argument("--force-balance-atol", type=float, default=None,
                    help=f"maximum accepted force_balance_relerr (default {FORCE_BALANCE_ATOL:g})")
    a = ap.parse_args(argv)
    try:
        s = validate_result(a.result_directory, force_balance_atol=a.force_balance_atol)
    except ValidationError as exc:
        print(f"INVALID: {exc}", file=sys.stderr)
        return 1
    except Exception as exc:  # a crash of the validator is never an acceptance
        print(f"INVALID: validator failure ({type(exc).__name__}): {exc}", file=sys.stderr)
        return 2
    print(f"VALID: case={s['case_name']} backend={s['backend']} precision={s['precision']} "
          f"step={s['completed_step']} history_rows={s['history_rows']} coverage={s['coverage']} "
          f"force_balance_max={s['force_balance_max']:.6e} atol={s['force_balance_atol']:.6e}")
    if s["force_balance_recomputed"] is not None:
        print(f"force_balance_recomputed={s['force_balance_recomputed']:.6e}")
    print(f"fingerprint={s['config_fingerprint']}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
