"""Case runner of the HIP path: PyIB's JSON case format and run artifacts.

Mirrors ``src/pyib/solver/ib_case.py``: the versioned JSON schema
(``DEFAULT_CONFIG`` 56-97, ``validate_config`` 262-441, unknown-key
rejection), the runtime build (633-809), the step loop with time anchored to
``step*dt`` (1517-1603), the 41-column ``history.csv`` (112-154, 1129-1224),
atomic ``checkpoint.npz`` schema 3 and ``final_field.npz`` (823-911),
validated resume (914-1121) and signal-safe stop.  The schema grows only
additively: ``runtime.device`` accepts ``"hip"``, ``runtime.precision``
accepts ``"mixed"``, ``mesh.kind`` accepts ``"nested_box"`` (finest box
half extents in diameters, each coarser level doubling), and
``solver.mixed_inner_tolerance`` sets the float32 residual reduction between
float64 reliable updates (default 0.1).
Errors keep the reference's types: ``CaseConfigError`` for configuration,
checkpoint and history problems, ``RuntimeError`` for solver and invariant
failures, ``FloatingPointError`` for non-finite diagnostics and
``FileExistsError`` for an occupied output directory.
"""
from __future__ import annotations

import copy
import csv
import hashlib
import json
import os
import platform
import signal
import time
import uuid
import zipfile
from dataclasses import dataclass
from datetime import datetime, timezone
from pathlib import Path
from typing import Any

import numpy as np
import torch

from . import mesh as M

SCHEMA_VERSION = 1
RESULT_SCHEMA_VERSION = 5
CHECKPOINT_SCHEMA_VERSION = 3
BACKEND = "hip/pyib_hip"

DEFAULT_CONFIG: dict[str, Any] = {
    "schema_version": SCHEMA_VERSION,
    "name": "fixed_sphere",
    "case": {"kind": "fixed_sphere"},
    "runtime": {"device": "cpu", "precision": "float64"},
    "mesh": {"kind": "uniform_box", "dx": 0.25, "bounds": [[-1.0, 1.0], [-1.0, 1.0], [-1.0, 1.0]],
             "verbose": False},
    "physics": {"reynolds": 100.0, "density": 1.0, "inflow_speed": 1.0},
    "time": {"dt": 0.01, "steps": 5, "inner_iterations": 4},
    "ib": {"radius": 0.25, "center": [0.0, 0.0, 0.0], "marker_spacing": 0.25,
           "force_mode": "accumulated_explicit", "implicit": False, "alpha": 1.0,
           "motion": {"kind": "none", "axis": [0.0, 1.0, 0.0], "amplitude": 0.0, "frequency": 1.0, "phase": 0.0}},
    "solver": {"sparse": True, "cg_check_interval": 20, "cg_max_iterations": 1000,
               "fail_on_cg_nonconvergence": False},
    "output": {"sample_every": 10, "checkpoint_every": 1000, "field_every": 0, "field_start_step": 0,
               "save_field": True},
}

_ALLOWED_KEYS = {
    "root": {"schema_version", "name", "case", "runtime", "mesh", "physics", "time", "ib", "solver", "output"},
    "case": {"kind"},
    "runtime": {"device", "precision"},
    "mesh": {"kind", "dx", "bounds", "cells_per_diameter", "verbose", "half_extents_d", "levels"},
    "physics": {"reynolds", "density", "inflow_speed"},
    "time": {"dt", "steps", "inner_iterations"},
    "ib": {"radius", "center", "marker_spacing", "force_mode", "implicit", "alpha", "motion"},
    "motion": {"kind", "axis", "amplitude", "frequency", "phase", "spin_ratio"},
    "solver": {"sparse", "cg_check_interval", "cg_max_iterations", "fail_on_cg_nonconvergence",
               "mixed_inner_tolerance"},
    "output": {"sample_every", "checkpoint_every", "field_every", "field_start_step", "save_field"},
}

HISTORY_COLUMNS = [
    "step", "time", "Cd", "Cy", "Cz", "body_force_x", "body_force_y", "body_force_z",
    "force_fluid_x", "force_fluid_y", "force_fluid_z", "body_moment_x", "body_moment_y", "body_moment_z",
    "moment_fluid_x", "moment_fluid_y", "moment_fluid_z", "interior_momentum_x", "interior_momentum_y",
    "interior_momentum_z", "interior_angular_momentum_x", "interior_angular_momentum_y",
    "interior_angular_momentum_z", "interior_momentum_rate_x", "interior_momentum_rate_y",
    "interior_momentum_rate_z", "interior_angular_momentum_rate_x", "interior_angular_momentum_rate_y",
    "interior_angular_momentum_rate_z", "interior_indicator_volume", "slip_max", "slip_rms", "speed_max", "cfl",
    "mass_residual_inf", "force_balance_relerr", "ib_force_updates", "helmholtz_iterations",
    "helmholtz_relative_residual", "helmholtz_tolerance_ratio", "milliseconds_per_step",
]
_INT_COLUMNS = {"step", "ib_force_updates", "helmholtz_iterations"}


class CaseConfigError(ValueError):
    """The case configuration is incomplete, inconsistent or unsupported."""


# ------------------------------------------------------------------ config
def _deep_merge(base, update):
    out = copy.deepcopy(base)
    for k, v in update.items():
        if isinstance(v, dict) and isinstance(out.get(k), dict):
            out[k] = _deep_merge(out[k], v)
        else:
            out[k] = copy.deepcopy(v)
    return out


def _reject_unknown(section, values):
    unknown = sorted(set(values).difference(_ALLOWED_KEYS[section]))
    if unknown:
        raise CaseConfigError(f"unknown {section} configuration keys: {unknown}")


def _positive(name, value):
    try:
        x = float(value)
    except (TypeError, ValueError) as exc:
        raise CaseConfigError(f"{name} must be numeric") from exc
    if not np.isfinite(x) or x <= 0.0:
        raise CaseConfigError(f"{name} must be finite and positive")
    return x


def _positive_int(name, value):
    if isinstance(value, bool):
        raise CaseConfigError(f"{name} must be a positive integer")
    try:
        x = int(value)
    except (TypeError, ValueError) as exc:
        raise CaseConfigError(f"{name} must be a positive integer") from exc
    if x <= 0 or x != float(value):
        raise CaseConfigError(f"{name} must be a positive integer")
    return x


def _nonnegative_int(name, value):
    if isinstance(value, bool):
        raise CaseConfigError(f"{name} must be a non-negative integer")
    try:
        x = int(value)
    except (TypeError, ValueError) as exc:
        raise CaseConfigError(f"{name} must be a non-negative integer") from exc
    try:
        exact = x == float(value)
    except (TypeError, ValueError):
        exact = False
    if x < 0 or not exact:
        raise CaseConfigError(f"{name} must be a non-negative integer")
    return x


def _boolean(name, value):
    if not isinstance(value, bool):
        raise CaseConfigError(f"{name} must be true or false")
    return value


def validate_config(raw: dict[str, Any]) -> dict[str, Any]:
    """ib_case.py:262-441 (+ additive hip/mixed/nested_box)."""
    if not isinstance(raw, dict):
        raise CaseConfigError("the case file must contain a JSON object")
    _reject_unknown("root", raw)
    for section in ("case", "runtime", "mesh", "physics", "time", "ib", "solver", "output"):
        if section in raw:
            if not isinstance(raw[section], dict):
                raise CaseConfigError(f"{section} must be a JSON object")
            _reject_unknown(section, raw[section])
    if isinstance(raw.get("ib"), dict) and "motion" in raw["ib"]:
        if not isinstance(raw["ib"]["motion"], dict):
            raise CaseConfigError("ib.motion must be a JSON object")
        _reject_unknown("motion", raw["ib"]["motion"])
    config = _deep_merge(DEFAULT_CONFIG, raw)
    if isinstance(config["schema_version"], bool):
        raise CaseConfigError("schema_version must be an integer")
    try:
        sv = int(config["schema_version"])
        exact = sv == float(config["schema_version"])
    except (TypeError, ValueError):
        raise CaseConfigError("schema_version must be an integer") from None
    if not exact or sv != SCHEMA_VERSION:
        raise CaseConfigError(f"unsupported schema_version={config['schema_version']}; expected {SCHEMA_VERSION}")
    config["schema_version"] = sv
    if not isinstance(config["name"], str) or not config["name"].strip():
        raise CaseConfigError("name must be a non-empty string")
    case_kind = str(config["case"]["kind"])
    if case_kind not in {"fixed_sphere", "oscillating_sphere", "rotating_sphere"}:
        raise CaseConfigError(f"unsupported case.kind={case_kind!r}")
    device = str(config["runtime"]["device"])
    precision = str(config["runtime"]["precision"])
    if device not in {"cpu", "cuda", "hip"}:
        raise CaseConfigError("runtime.device must be 'cpu', 'cuda' or 'hip'")
    if precision not in {"float64", "float32", "mixed"}:
        raise CaseConfigError("runtime.precision must be 'float64', 'float32' or 'mixed'")
    if precision == "mixed" and device != "hip":
        raise CaseConfigError("runtime.precision='mixed' requires runtime.device='hip'")
    mesh = config["mesh"]
    mesh_kind = str(mesh["kind"])
    if mesh_kind not in {"uniform_box", "sphere_wake", "sphere_concentric", "nested_box"}:
        raise CaseConfigError(f"unsupported mesh.kind={mesh_kind!r}")
    if mesh_kind != "nested_box" and ("half_extents_d" in mesh or "levels" in mesh):
        raise CaseConfigError("mesh.half_extents_d/levels are only valid for mesh.kind='nested_box'")
    if mesh_kind == "uniform_box":
        mesh["dx"] = _positive("mesh.dx", mesh["dx"])
        bounds = np.asarray(mesh["bounds"], dtype=float)
        if bounds.shape != (3, 2) or not np.all(np.isfinite(bounds)):
            raise CaseConfigError("mesh.bounds must be three finite [lower, upper] pairs")
        if np.any(bounds[:, 1] <= bounds[:, 0]):
            raise CaseConfigError("every mesh bound must have upper > lower")
        cells = (bounds[:, 1] - bounds[:, 0]) / mesh["dx"]
        if not np.allclose(cells, np.rint(cells), rtol=0.0, atol=1.0e-10):
            raise CaseConfigError("uniform-box side lengths must be integer multiples of mesh.dx")
        mesh["bounds"] = bounds.tolist()
    else:
        mesh["cells_per_diameter"] = _positive_int("mesh.cells_per_diameter", mesh.get("cells_per_diameter", 16))
        if mesh["cells_per_diameter"] % 8:
            raise CaseConfigError("nested-sphere cells_per_diameter must be a multiple of 8")
        mesh.pop("dx", None)
        mesh.pop("bounds", None)
        if mesh_kind == "nested_box":
            he = np.asarray(mesh.get("half_extents_d", [3.0, 2.5, 2.5]), dtype=float)
            if he.shape != (3,) or not np.all(np.isfinite(he)) or np.any(he <= 0):
                raise CaseConfigError("mesh.half_extents_d must be three positive numbers")
            n2 = he * mesh["cells_per_diameter"] * 2.0
            if not np.allclose(n2, np.rint(n2), atol=1e-9) or np.any(np.rint(n2).astype(int) % 4):
                raise CaseConfigError("mesh.half_extents_d must give finest cell counts divisible by 4")
            mesh["half_extents_d"] = he.tolist()
            mesh["levels"] = _positive_int("mesh.levels", mesh.get("levels", 3))
    physics = config["physics"]
    physics["reynolds"] = _positive("physics.reynolds", physics["reynolds"])
    physics["density"] = _positive("physics.density", physics["density"])
    physics["inflow_speed"] = float(physics["inflow_speed"])
    if not np.isfinite(physics["inflow_speed"]):
        raise CaseConfigError("physics.inflow_speed must be finite")
    tc = config["time"]
    tc["dt"] = _positive("time.dt", tc["dt"])
    tc["steps"] = _positive_int("time.steps", tc["steps"])
    tc["inner_iterations"] = _positive_int("time.inner_iterations", tc["inner_iterations"])
    ib = config["ib"]
    ib["radius"] = _positive("ib.radius", ib["radius"])
    centre = np.asarray(ib["center"], dtype=float)
    if centre.shape != (3,) or not np.all(np.isfinite(centre)):
        raise CaseConfigError("ib.center must contain three finite coordinates")
    ib["center"] = centre.tolist()
    ib["marker_spacing"] = _positive("ib.marker_spacing", ib["marker_spacing"])
    ib["alpha"] = _positive("ib.alpha", ib["alpha"])
    if ib["alpha"] > 1.0:
        raise CaseConfigError("ib.alpha must be <= 1.0")
    force_mode = str(ib["force_mode"])
    if force_mode not in {"direct", "accumulated_explicit"}:
        raise CaseConfigError("ib.force_mode must be 'accumulated_explicit' or 'direct'")
    ib["force_mode"] = force_mode
    ib["implicit"] = _boolean("ib.implicit", ib["implicit"])
    if force_mode == "accumulated_explicit":
        if ib["implicit"]:
            raise CaseConfigError("accumulated_explicit is an explicit IB method")
        if ib["alpha"] != 1.0:
            raise CaseConfigError("accumulated_explicit requires ib.alpha=1.0")
        if tc["inner_iterations"] < 2:
            raise CaseConfigError("accumulated_explicit requires at least two fluid inner iterations")
    elif ib["implicit"] and ib["alpha"] > 0.7:
        raise CaseConfigError("implicit IB currently requires ib.alpha <= 0.7")
    motion = ib["motion"]
    motion_kind = str(motion["kind"])
    expected = {"fixed_sphere": "none", "oscillating_sphere": "oscillation", "rotating_sphere": "rotation"}[case_kind]
    if motion_kind != expected:
        raise CaseConfigError(f"case.kind={case_kind!r} requires ib.motion.kind={expected!r}")
    axis = np.asarray(motion["axis"], dtype=float)
    if axis.shape != (3,) or not np.all(np.isfinite(axis)) or np.linalg.norm(axis) == 0.0:
        raise CaseConfigError("ib.motion.axis must be a finite non-zero vector")
    motion["axis"] = (axis / np.linalg.norm(axis)).tolist()
    motion["amplitude"] = float(motion["amplitude"])
    motion["frequency"] = _positive("ib.motion.frequency", motion["frequency"])
    motion["phase"] = float(motion["phase"])
    if not np.isfinite(motion["amplitude"]) or motion["amplitude"] < 0.0:
        raise CaseConfigError("ib.motion.amplitude must be finite and non-negative")
    if not np.isfinite(motion["phase"]):
        raise CaseConfigError("ib.motion.phase must be finite")
    if motion_kind == "rotation":
        if "spin_ratio" not in motion:
            raise CaseConfigError("rotating_sphere requires ib.motion.spin_ratio")
        motion["spin_ratio"] = float(motion["spin_ratio"])
        if not np.isfinite(motion["spin_ratio"]) or motion["spin_ratio"] == 0.0:
            raise CaseConfigError("ib.motion.spin_ratio must be finite and non-zero")
    elif "spin_ratio" in motion:
        raise CaseConfigError("ib.motion.spin_ratio is only valid for rotating_sphere")
    if case_kind == "oscillating_sphere" and motion["amplitude"] == 0.0:
        raise CaseConfigError("oscillating_sphere requires a positive motion amplitude")
    if case_kind == "fixed_sphere" and motion["amplitude"] != 0.0:
        raise CaseConfigError("fixed_sphere requires zero motion amplitude")
    if case_kind == "rotating_sphere" and motion["amplitude"] != 0.0:
        raise CaseConfigError("rotating_sphere requires zero translational amplitude")
    if case_kind in {"fixed_sphere", "rotating_sphere"} and physics["inflow_speed"] == 0.0:
        raise CaseConfigError(f"{case_kind} requires non-zero physics.inflow_speed")
    out = config["output"]
    out["sample_every"] = _positive_int("output.sample_every", out["sample_every"])
    out["checkpoint_every"] = _nonnegative_int("output.checkpoint_every", out["checkpoint_every"])
    out["save_field"] = _boolean("output.save_field", out["save_field"])
    out["field_every"] = _nonnegative_int("output.field_every", out["field_every"])
    out["field_start_step"] = _nonnegative_int("output.field_start_step", out["field_start_step"])
    config["mesh"]["verbose"] = _boolean("mesh.verbose", config["mesh"].get("verbose", False))
    sv = config["solver"]
    sv["sparse"] = _boolean("solver.sparse", sv["sparse"])
    sv["cg_check_interval"] = _positive_int("solver.cg_check_interval", sv["cg_check_interval"])
    sv["cg_max_iterations"] = _positive_int("solver.cg_max_iterations", sv["cg_max_iterations"])
    sv["fail_on_cg_nonconvergence"] = _boolean("solver.fail_on_cg_nonconvergence", sv["fail_on_cg_nonconvergence"])
    if "mixed_inner_tolerance" in sv:
        sv["mixed_inner_tolerance"] = _positive("solver.mixed_inner_tolerance", sv["mixed_inner_tolerance"])
        if sv["mixed_inner_tolerance"] >= 1.0:
            raise CaseConfigError("solver.mixed_inner_tolerance must be < 1")
    if mesh_kind in {"sphere_wake", "sphere_concentric", "nested_box"}:
        if not np.allclose(centre, 0.0) or not np.isclose(ib["radius"], 0.25):
            raise CaseConfigError("nested sphere meshes require ib.center=[0,0,0] and ib.radius=0.25")
    return config


def load_case_config(path):
    src = Path(path)
    try:
        raw = json.loads(src.read_text(encoding="utf-8"))
    except OSError as exc:
        raise CaseConfigError(f"cannot read case file {src}: {exc}") from exc
    except json.JSONDecodeError as exc:
        raise CaseConfigError(f"invalid JSON in {src}: {exc}") from exc
    return validate_config(raw)


def _immutable_config(config):
    c = copy.deepcopy(config)
    c.pop("name", None)
    c.pop("output", None)
    c["time"].pop("steps", None)
    return c


def config_fingerprint(config):
    """ib_case.py:455-467."""
    payload = json.dumps(_immutable_config(config), sort_keys=True, separators=(",", ":"), allow_nan=False)
    return hashlib.sha256(payload.encode("utf-8")).hexdigest()


_LAYOUT_NAMES = ("nodes", "cells", "boundary_faces", "boundary_fam", "pos_ref", "dx_uniform", "index_cell",
                 "faces_to_nodes", "faces_to_cells", "edges_to_nodes", "faces_to_edges", "bc_node", "bc_edge",
                 "bc_face", "bc_cell", "interface_pairs", "link_fe_insert")


def layout_fingerprint(mesh, markers, widths):
    """ib_case.py:490-510: sha256 of topology, ordering, boundary labels and quadrature."""
    d = hashlib.sha256()
    for name in _LAYOUT_NAMES:
        a = np.ascontiguousarray(np.asarray(getattr(mesh, name)))
        d.update(name.encode("utf-8") + b"\0")
        d.update(a.dtype.str.encode("ascii") + b"\0")
        d.update(json.dumps(a.shape).encode("ascii") + b"\0")
        d.update(a.tobytes(order="C"))
    for name, v in (("markers", markers), ("marker_widths", widths)):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        d.update(name.encode("utf-8") + b"\0")
        d.update(json.dumps(a.shape).encode("ascii") + b"\0")
        d.update(a.tobytes(order="C"))
    return d.hexdigest()


_SOURCE_FILES = ("case.py", "core.py", "driver.py", "geometry.py", "grid.py", "immersed.py", "interior.py",
                 "kinematics.py", "layout.py", "mesh.py", "tables.py", "csrc/pyib_hip.hip", "csrc/grid.hip",
                 "csrc/geometry.cpp", "csrc/mesh.cpp", "csrc/layout.cpp", "csrc/implicit.hip")


def source_hashes():
    here = Path(__file__).resolve().parent
    return {name: hashlib.sha256((here / name).read_bytes()).hexdigest() for name in _SOURCE_FILES}


def fibonacci_sphere(count, radius, centre):
    """ib_case.py:554-563."""
    if count < 4:
        raise ValueError("a sphere surface requires at least four markers")
    i = np.arange(count, dtype=float)
    phi = np.arccos(1.0 - 2.0 * (i + 0.5) / count)
    theta = np.pi * (1.0 + np.sqrt(5.0)) * i
    unit = np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)], axis=1)
    return np.asarray(centre, dtype=float) + radius * unit


def sphere_center(config, t):
    """ib_case.py:587-603."""
    c = np.asarray(config["ib"]["center"], dtype=float)
    mo = config["ib"]["motion"]
    if mo["kind"] != "oscillation":
        return c
    ang = 2.0 * np.pi * float(mo["frequency"]) * float(t) + float(mo["phase"])
    return c + float(mo["amplitude"]) * np.sin(ang) * np.asarray(mo["axis"], dtype=float)


def rotating_sphere_velocity(markers, centre, axis, spin_ratio, radius, reference_speed):
    """ib_case.py:566-584."""
    pts = np.asarray(markers, dtype=float)
    direction = np.asarray(axis, dtype=float)
    omega = direction / np.linalg.norm(direction) * float(spin_ratio) * float(reference_speed) / float(radius)
    return np.cross(omega, pts - np.asarray(centre, dtype=float))


def _validate_ib_support(config, dx, bounds):
    """ib_case.py:614-630."""
    ib = config["ib"]
    centre = np.asarray(ib["center"], dtype=float)
    radius = float(ib["radius"])
    mo = ib["motion"]
    exc = np.zeros(3)
    if mo["kind"] == "oscillation":
        exc = abs(float(mo["amplitude"])) * np.abs(np.asarray(mo["axis"]))
    clearance = radius + 2.5 * dx
    lo, hi = centre - exc - clearance, centre + exc + clearance
    if np.any(lo < bounds[:, 0] - 1.0e-12) or np.any(hi > bounds[:, 1] + 1.0e-12):
        raise CaseConfigError("the sphere or its IB support leaves the finest uniform mesh region; "
                              "increase the domain/refined region, reduce radius/amplitude, or reduce dx")


def build_mesh(config):
    """Mesh of a validated config -> (mesh, finest dx, finest bounds, summary)."""
    mc = config["mesh"]
    radius = float(config["ib"]["radius"])
    kind = mc["kind"]
    if kind == "uniform_box":
        dx = float(mc["dx"])
        bounds = np.asarray(mc["bounds"], dtype=float)
        _validate_ib_support(config, dx, bounds)
        mesh = M.make_nested_mesh(dx, M.uniform_boxes(bounds), verbose=bool(mc["verbose"]))
        return mesh, dx, mesh_summary(config, mesh)
    ddx = int(mc["cells_per_diameter"])
    dx = 2.0 * radius / ddx
    if kind == "sphere_wake":
        b = np.asarray(M.WAKE_BOUNDS_D[0], dtype=float) * 2.0 * radius
        _validate_ib_support(config, dx, b.reshape(3, 2))
        boxes = M.wake_boxes()
    elif kind == "sphere_concentric":
        hw = float(M.CONCENTRIC_HALF_WIDTHS_D[0]) * 2.0 * radius
        _validate_ib_support(config, dx, np.asarray([[-hw, hw]] * 3))
        boxes = M.concentric_boxes()
    else:
        he = np.asarray(mc["half_extents_d"], dtype=float) * 2.0 * radius
        _validate_ib_support(config, dx, np.stack([-he, he], axis=1))
        boxes = M.nested_box_boxes(mc["half_extents_d"], int(mc["levels"]))
    from . import setup_cache
    mesh = setup_cache.cached("mesh", _mesh_params(config),
                              lambda: M.make_nested_mesh(M.DIAMETER / ddx, boxes, verbose=bool(mc["verbose"])),
                              M.CombinedMesh)
    return mesh, dx, mesh_summary(config, mesh)


def mesh_summary(config, mesh):
    """metadata.json's ``mesh`` record (before the face/edge/marker counts are added)."""
    mc = config["mesh"]
    if mc["kind"] == "uniform_box":
        return {"kind": "uniform_box", "dx": float(mc["dx"]), "bounds": np.asarray(mc["bounds"], float).tolist(),
                "cells": int(mesh.cells.shape[0])}
    ddx = int(mc["cells_per_diameter"])
    return {"kind": mc["kind"], "diameter": M.DIAMETER, "D_over_dx": ddx, "dx_fine": M.DIAMETER / ddx,
            "layer_cells": [int(v) for v in mesh.layer_cell_counts], "total_cells": int(mesh.cells.shape[0]),
            "nodes": int(mesh.nodes.shape[0]), "layer_shape": [list(s) for s in mesh.layer_shape]}


def _mesh_params(config):
    mc = config["mesh"]
    return {k: mc.get(k) for k in ("kind", "cells_per_diameter", "half_extents_d", "levels", "dx", "bounds")}


_SETUP_MEMO = {}


def setup_objects(config, mesh, log=None):
    """Geometry, gather tables and structured layout of a mesh (cached on disk when
    PYIB_HIP_SETUP_CACHE is set, and in this process for the mesh object last seen).
    ``log`` (a callable taking one string) is told when each stage ends."""
    key = (id(mesh), json.dumps(_mesh_params(config), sort_keys=True))
    hit = _SETUP_MEMO.get(key)
    if hit is not None and hit[0] is mesh:
        return hit[1]
    out = _setup_objects(config, mesh, log or (lambda m: None))
    _SETUP_MEMO.clear()
    _SETUP_MEMO[key] = (mesh, out)
    return out


def _setup_objects(config, mesh, log):
    from . import setup_cache
    from .geometry import Geometry
    from .layout import Composite, Level, StructuredLayout, build_layout
    from .tables import Tables, build_tables
    params = _mesh_params(config)
    if config["mesh"]["kind"] == "uniform_box":
        params["fingerprint"] = hashlib.sha256(np.ascontiguousarray(mesh.nodes).tobytes()).hexdigest()
    t0 = time.perf_counter()
    geo = setup_cache.cached("geometry", params, lambda: Geometry(mesh), Geometry)
    log(f"geometry ready ({time.perf_counter() - t0:.1f} s)")
    tables = setup_cache.cached("tables", params, lambda: build_tables(geo), Tables)
    log(f"gather tables ready ({time.perf_counter() - t0:.1f} s)")
    lay = setup_cache.cached("layout", dict(params, layout_rules="per-pass-regularity-3"),
                             lambda: build_layout(mesh, geo, tables), StructuredLayout,
                             {"Composite": Composite, "Level": Level})
    log(f"structured layout ready ({time.perf_counter() - t0:.1f} s)")
    return geo, tables, lay


def uses_cavity(config):
    """Quiescent-fluid cases use the no-slip box (ib_case.py:717-718 for sphere_concentric)."""
    kind = config["mesh"]["kind"]
    return kind == "sphere_concentric" or (kind == "nested_box" and float(config["physics"]["inflow_speed"]) == 0.0)


# ----------------------------------------------------------------- runtime
@dataclass
class CaseRuntime:
    config: dict
    mesh: Any
    mesh_summary: dict
    core: Any
    driver: Any
    ib: Any
    load_correction: Any
    reference_markers: np.ndarray
    marker_widths: np.ndarray
    reference_force: float
    reference_speed: float
    dx: float
    dt: float
    steps: int
    inner_iterations: int
    layout_fingerprint: str
    source_hashes: dict


def build_runtime(config, mesh=None, device="cuda", log=None):
    """ib_case.py:633-809 on the HIP path.  ``log`` (a callable taking one string) is told
    when each set-up stage ends; a 40-cells-per-diameter mesh takes minutes to set up."""
    from .core import HipCore
    from .driver import HipDriver
    from .immersed import HipImmersedBoundary
    from .interior import HipLoadCorrection
    from .kinematics import Oscillation

    config = validate_config(config)
    ibc = config["ib"]
    radius = float(ibc["radius"])
    centre = np.asarray(ibc["center"], dtype=float)
    say = log or (lambda m: None)
    if mesh is None:
        mesh, dx, summary = build_mesh(config)
        say(f"mesh ready ({summary.get('total_cells', summary.get('cells'))} cells)")
    else:
        dx = float(np.asarray(mesh.dx_uniform)[0])
        summary = mesh_summary(config, mesh)
    precision = config["runtime"]["precision"]
    if precision == "float32":
        raise CaseConfigError("the HIP path runs float64 or mixed precision")
    geo, tables, lay = setup_objects(config, mesh, log)
    core = HipCore(mesh, device=device, precisions=("float64", "float32") if precision == "mixed" else ("float64",),
                   geometry=geo, tables=tables)
    say("device operators uploaded")
    ph = config["physics"]
    U, Re, rho = float(ph["inflow_speed"]), float(ph["reynolds"]), float(ph["density"])
    mo = ibc["motion"]
    if U != 0.0:
        nu, uref = abs(U) * 2.0 * radius / Re, abs(U)
    elif mo["kind"] == "oscillation":
        uref = 2.0 * np.pi * float(mo["frequency"]) * float(mo["amplitude"])
        nu = uref * 2.0 * radius / Re
    else:
        raise CaseConfigError("a stationary sphere requires non-zero physics.inflow_speed")
    sv = config["solver"]
    drv = HipDriver(core, mesh, visc=rho * nu, rou=rho, cg_max_iterations=int(sv["cg_max_iterations"]),
                    fail_on_cg_nonconvergence=bool(sv["fail_on_cg_nonconvergence"]),
                    cg_check_interval=int(sv["cg_check_interval"]), precision=precision,
                    mixed_inner_tol=float(sv.get("mixed_inner_tolerance", 0.1)), layout=lay)
    say("structured grid uploaded")
    if uses_cavity(config):
        drv.set_cavity(U_lid=0.0, spanwise_zerograd=False)
    dt = float(config["time"]["dt"])
    count = max(12, int(round(4.0 * np.pi * radius ** 2 / float(ibc["marker_spacing"]) ** 2)))
    markers = fibonacci_sphere(count, radius, centre)
    widths = np.full(count, (4.0 * np.pi * radius ** 2 * dx / count) ** (1.0 / 3.0))
    lfp = layout_fingerprint(mesh, markers, widths)
    ib = HipImmersedBoundary(mesh, device=device)
    mode = "accumulated_explicit" if ibc["force_mode"] == "accumulated_explicit" else "replacement"
    if mo["kind"] == "oscillation":
        motion = Oscillation(markers, axis=mo["axis"], A=float(mo["amplitude"]), freq=float(mo["frequency"]),
                             phase=float(mo["phase"]))
        m0, v0 = motion(0.0)
        drv.set_ib(ib, m0, widths, v0, force_mode=mode)
        drv.set_motion(motion, t0=0.0)
    elif mo["kind"] == "rotation":
        vel = rotating_sphere_velocity(markers, centre, mo["axis"], float(mo["spin_ratio"]), radius, uref)
        drv.set_ib(ib, markers, widths, vel, force_mode=mode)
    else:
        drv.set_ib(ib, markers, widths, np.zeros((count, 3)), force_mode=mode)
    drv.ib_alpha = float(ibc["alpha"])
    drv.ib_implicit = bool(ibc["implicit"])      # ib_case.py:766
    mt = float(mo["amplitude"]) * np.abs(np.asarray(mo["axis"], dtype=float)) if mo["kind"] == "oscillation" \
        else np.zeros(3)
    load = HipLoadCorrection(core.cpos[:core.ncell], 1.0 / core.vol_ci, reference_center=centre,
                             maximum_translation=mt, radius=radius, smoothing_width=dx, density=rho,
                             force_support_width=1.5 * ib.hh, device=device)
    fref = 0.5 * rho * uref ** 2 * (np.pi * radius ** 2)
    summary.update({"faces": int(core.nface), "edges": int(core.nedge), "markers": count})
    return CaseRuntime(config, mesh, summary, core, drv, ib, load, markers, widths, fref, uref, dx, dt,
                       int(config["time"]["steps"]), int(config["time"]["inner_iterations"]), lfp, source_hashes())


def initial_state(rt):
    """ib_case.py:812-820."""
    core = rt.core
    s0 = float(rt.config["physics"]["inflow_speed"]) * core.epos[:, 1] * core.te[:, 2]
    s, _, vc = rt.driver.calc_vel(core.tensor(s0))
    rt.load_correction.initialize(vc, sphere_center(rt.config, rt.driver.time))
    return s, s.clone(), vc


def diagnostics(rt, step, s, vc, milliseconds_per_step):
    """ib_case.py:1129-1224."""
    import ctypes as C
    from . import _native as N
    from .core import dev_stream, ptr
    core, drv = rt.core, rt.driver
    load = rt.load_correction.latest_load
    if load is None:
        raise RuntimeError("closed-body load correction was not advanced for this step")
    body_force = np.asarray(load.body_force)
    coeff = body_force / rt.reference_force
    drv.ib.stencil(drv.ib_markers, drv._marker_version)
    mv = drv.ib.interp(vc).cpu().numpy()
    slip = np.linalg.norm(mv - drv.ib_vel.cpu().numpy(), axis=1)
    U = drv.last_face_flux
    if U is None:
        _, U, _ = drv.calc_vel(s)
    tmp = torch.empty(core.ncell + 2 * N.MAX_PARTIALS, dtype=torch.float64, device=core.device)
    out2 = torch.zeros(2, dtype=torch.float64, device=core.device)
    N.call("pyib_diag_max", core.struct(), ptr(U), ptr(vc), ptr(tmp), ptr(out2), dev_stream())
    mass_max, speed_max = (float(v) for v in out2.cpu().numpy())
    lag = (drv.ib_force_L.cpu().numpy() * (drv.ib_ds ** 3)[:, None]).sum(axis=0)
    ff = np.asarray(load.force_fluid)
    den = max(float(np.linalg.norm(lag)), float(np.linalg.norm(ff)), 1.0e-300)
    balance = float(np.linalg.norm(ff - lag)) / den
    row = {"step": int(step), "time": float(drv.time), "Cd": float(coeff[0]), "Cy": float(coeff[1]),
           "Cz": float(coeff[2])}
    for name, v in (("body_force", body_force), ("force_fluid", ff), ("body_moment", load.body_moment),
                    ("moment_fluid", load.moment_fluid),
                    ("interior_momentum", load.interior_momenta.linear_momentum),
                    ("interior_angular_momentum", load.interior_momenta.angular_momentum),
                    ("interior_momentum_rate", load.linear_momentum_rate),
                    ("interior_angular_momentum_rate", load.angular_momentum_rate)):
        for k, ax in enumerate("xyz"):
            row[f"{name}_{ax}"] = float(v[k])
    row.update({
        "interior_indicator_volume": float(load.interior_momenta.indicator_volume),
        "slip_max": float(np.max(slip)), "slip_rms": float(np.sqrt(np.mean(slip ** 2))),
        "speed_max": speed_max, "cfl": speed_max * rt.dt / rt.dx, "mass_residual_inf": mass_max,
        "force_balance_relerr": balance, "ib_force_updates": int(drv.ib_force_updates_last_step),
        "helmholtz_iterations": int(drv.helmholtz_iterations_last_step),
        "helmholtz_relative_residual": float(drv.helmholtz_relative_residual_last_step),
        "helmholtz_tolerance_ratio": float(drv.helmholtz_tolerance_ratio_last_step),
        "milliseconds_per_step": float(milliseconds_per_step),
    })
    numeric = np.asarray([row[n] for n in HISTORY_COLUMNS], dtype=float)
    if not np.all(np.isfinite(numeric)):
        raise FloatingPointError(f"non-finite diagnostic at step {step}: {row}")
    if balance > 1.0e-9:
        raise RuntimeError(f"IB force-spread balance failed at step {step}: {balance:.3e}")
    expected = 1 if rt.config["ib"]["force_mode"] == "accumulated_explicit" else rt.inner_iterations
    if row["ib_force_updates"] != expected:
        raise RuntimeError(f"IB update-count invariant failed at step {step}: {row['ib_force_updates']} != {expected}")
    return row


# --------------------------------------------------------------- artifacts
def _fsync_parent(path):
    try:
        fd = os.open(path.parent, os.O_RDONLY)
    except OSError:
        return
    try:
        os.fsync(fd)
    except OSError:
        pass
    finally:
        os.close(fd)


def _atomic_json(path, payload):
    tmp = path.with_name(f".{path.name}.{uuid.uuid4().hex}.tmp")
    try:
        with tmp.open("x", encoding="utf-8") as h:
            h.write(json.dumps(payload, indent=2, allow_nan=False) + "\n")
            h.flush()
            os.fsync(h.fileno())
        os.replace(tmp, path)
        _fsync_parent(path)
    finally:
        if tmp.exists():
            tmp.unlink()


def _atomic_npz(path, **arrays):
    tmp = path.with_name(f".{path.stem}.{uuid.uuid4().hex}.tmp.npz")
    try:
        np.savez(tmp, **arrays)
        with tmp.open("r+b") as h:
            h.flush()
            os.fsync(h.fileno())
        os.replace(tmp, path)
        _fsync_parent(path)
    finally:
        if tmp.exists():
            tmp.unlink()


def _precision(rt):
    return rt.config["runtime"]["precision"]


def _checkpoint_payload(rt, run_id, step, s, s_old):
    """ib_case.py:823-870."""
    drv = rt.driver
    st = drv.ib_state_dict()
    load = rt.load_correction.latest_load
    if load is None or rt.load_correction.current is None:
        raise RuntimeError("cannot checkpoint before closed-body load bookkeeping")
    return {
        "checkpoint_schema": np.int64(CHECKPOINT_SCHEMA_VERSION), "update_phase": np.str_("post_step"),
        "run_id": np.str_(run_id), "config_fingerprint": np.str_(config_fingerprint(rt.config)),
        "layout_fingerprint": np.str_(rt.layout_fingerprint), "step": np.int64(step),
        "physical_time": np.float64(drv.time), "dt": np.float64(rt.dt), "previous_dt": np.float64(rt.dt),
        "precision": np.str_(_precision(rt)), "device": np.str_("hip"),
        "source_hashes_json": np.str_(json.dumps(rt.source_hashes, sort_keys=True, separators=(",", ":"))),
        "s": s.cpu().numpy(), "s_old": s_old.cpu().numpy(), "markers": drv.ib_markers.cpu().numpy(),
        "marker_velocity": drv.ib_vel.cpu().numpy(), "ib_force_mode": np.str_(rt.config["ib"]["force_mode"]),
        "ib_force_E": st["force_E"], "ib_force_L": st["force_L"],
        "ib_force_update_count": np.int64(st["force_update_count"]),
        "force_fluid": np.asarray(load.force_fluid), "moment_fluid": np.asarray(load.moment_fluid),
        "body_force": np.asarray(load.body_force), "body_moment": np.asarray(load.body_moment),
        "interior_momentum": np.asarray(load.interior_momenta.linear_momentum),
        "interior_angular_momentum": np.asarray(load.interior_momenta.angular_momentum),
        "interior_indicator_volume": np.asarray(load.interior_momenta.indicator_volume),
        "interior_momentum_rate": np.asarray(load.linear_momentum_rate),
        "interior_angular_momentum_rate": np.asarray(load.angular_momentum_rate),
    }


def save_checkpoint(path, rt, run_id, step, s, s_old):
    _atomic_npz(path, **_checkpoint_payload(rt, run_id, step, s, s_old))


def save_field_snapshot(path, rt, run_id, step, s, s_old, vc):
    """ib_case.py:879-911."""
    drv, core = rt.driver, rt.core
    if drv.last_momentum_residual is None or drv.last_step_dt != rt.dt:
        raise RuntimeError("no current-step momentum residual for pressure recovery")
    p = _checkpoint_payload(rt, run_id, step, s, s_old)
    for name in ("checkpoint_schema", "update_phase", "physical_time", "previous_dt"):
        p.pop(name)
    p.update(result_schema=np.int64(RESULT_SCHEMA_VERSION), time=np.float64(drv.time),
             vc=vc[:core.ncell].cpu().numpy(), cpos=core.cpos, cell_volume=1.0 / core.vol_ci,
             marker_volume_weight=drv.ib_ds ** 3,
             momentum_residual=drv.last_momentum_residual[:core.ncell].cpu().numpy(),
             helmholtz_relative_residual=np.float64(drv.helmholtz_relative_residual_last_step),
             helmholtz_tolerance_ratio=np.float64(drv.helmholtz_tolerance_ratio_last_step),
             face_cells=core.F2C, face_area=core.Af, index_cell=rt.ib.index_cell, lattice_origin=rt.ib.pos_ref,
             lattice_spacing=np.float64(rt.dx), sphere_center=np.asarray(sphere_center(rt.config, drv.time)),
             sphere_radius=np.float64(rt.config["ib"]["radius"]), density=np.float64(drv.rou),
             kinematic_viscosity=np.float64(drv.visc / drv.rou))
    _atomic_npz(path, **p)


def load_checkpoint(path, rt):
    """ib_case.py:914-1121 (device 'hip'; precision float64 or mixed)."""
    try:
        ck = np.load(path, allow_pickle=False)
    except (OSError, ValueError, EOFError, zipfile.BadZipFile) as exc:
        raise CaseConfigError(f"cannot read checkpoint {path}: {exc}") from exc
    drv, core = rt.driver, rt.core
    with ck:
        required = {"checkpoint_schema", "update_phase", "run_id", "config_fingerprint", "layout_fingerprint", "step",
                    "physical_time", "dt", "previous_dt", "precision", "device", "source_hashes_json", "s", "s_old",
                    "markers", "marker_velocity", "ib_force_mode", "ib_force_E", "ib_force_L",
                    "ib_force_update_count", "force_fluid", "moment_fluid", "body_force", "body_moment",
                    "interior_momentum", "interior_angular_momentum", "interior_indicator_volume",
                    "interior_momentum_rate", "interior_angular_momentum_rate"}
        missing = sorted(required.difference(ck.files))
        if missing:
            raise CaseConfigError(f"checkpoint is missing fields: {missing}")
        if int(ck["checkpoint_schema"]) != CHECKPOINT_SCHEMA_VERSION:
            raise CaseConfigError(f"unsupported checkpoint schema {int(ck['checkpoint_schema'])}; "
                                  f"expected {CHECKPOINT_SCHEMA_VERSION}")
        if str(np.asarray(ck["update_phase"]).item()) != "post_step":
            raise CaseConfigError("unsupported checkpoint update phase")
        if str(np.asarray(ck["config_fingerprint"]).item()) != config_fingerprint(rt.config):
            raise CaseConfigError("checkpoint configuration does not match the requested case")
        if str(np.asarray(ck["layout_fingerprint"]).item()) != rt.layout_fingerprint:
            raise CaseConfigError("checkpoint mesh/marker layout does not match this case")
        if float(ck["dt"]) != rt.dt:
            raise CaseConfigError("checkpoint dt does not match the requested case")
        if float(ck["previous_dt"]) != rt.dt:
            raise CaseConfigError("checkpoint previous_dt does not match the constant-step case")
        if str(np.asarray(ck["precision"]).item()) != _precision(rt):
            raise CaseConfigError("checkpoint precision does not match the requested case")
        if str(np.asarray(ck["device"]).item()) != "hip":
            raise CaseConfigError("checkpoint device does not match the requested case")
        try:
            old_hashes = json.loads(str(np.asarray(ck["source_hashes_json"]).item()))
        except (json.JSONDecodeError, TypeError) as exc:
            raise CaseConfigError("checkpoint source identity is invalid") from exc
        if old_hashes != rt.source_hashes:
            raise CaseConfigError("checkpoint numerical source identity does not match this solver")
        run_id = str(np.asarray(ck["run_id"]).item())
        try:
            uuid.UUID(run_id)
        except (ValueError, AttributeError) as exc:
            raise CaseConfigError("checkpoint run_id is invalid") from exc
        step = int(ck["step"])
        t = float(ck["physical_time"])
        expected_time = step * rt.dt
        tol = max(1.0e-12, 16.0 * np.finfo(np.float64).eps * max(abs(expected_time), rt.dt) * max(step, 1))
        if step < 0 or abs(t - expected_time) > tol:
            raise CaseConfigError("checkpoint step and physical time are inconsistent")
        s_raw, so_raw = np.asarray(ck["s"]), np.asarray(ck["s_old"])
        if s_raw.dtype.kind != "f" or so_raw.dtype.kind != "f":
            raise CaseConfigError("checkpoint flow fields must use floating-point arrays")
        if s_raw.shape != (core.nedge,) or so_raw.shape != (core.nedge,):
            raise CaseConfigError("checkpoint field shapes do not match the mesh")
        markers, mvel = np.asarray(ck["markers"]), np.asarray(ck["marker_velocity"])
        shape = tuple(drv.ib_markers.shape)
        if markers.shape != shape or mvel.shape != shape:
            raise CaseConfigError("checkpoint marker shapes do not match the case")
        fE, fL = np.asarray(ck["ib_force_E"]), np.asarray(ck["ib_force_L"])
        if fE.dtype.kind != "f" or fL.dtype.kind != "f":
            raise CaseConfigError("checkpoint IB forces must use floating-point arrays")
        if fE.shape != (core.ncell, 3) or fL.shape != shape:
            raise CaseConfigError("checkpoint IB-force shapes do not match the case")
        vec = {n: np.asarray(ck[n]) for n in ("force_fluid", "moment_fluid", "body_force", "body_moment",
                                               "interior_momentum", "interior_angular_momentum",
                                               "interior_momentum_rate", "interior_angular_momentum_rate")}
        if any(v.dtype.kind != "f" or v.shape != (3,) for v in vec.values()):
            raise CaseConfigError("checkpoint closed-body load vectors are invalid")
        iv = np.asarray(ck["interior_indicator_volume"])
        if iv.dtype.kind != "f" or iv.shape != ():
            raise CaseConfigError("checkpoint interior indicator volume is invalid")
        if not all(np.all(np.isfinite(a)) for a in (s_raw, so_raw, markers, mvel, fE, fL, iv, *vec.values())):
            raise CaseConfigError("checkpoint contains non-finite state")
        ltol = 1.0e-11
        positions, volume = core.cpos[:core.ncell], 1.0 / core.vol_ci
        iforce = (fE * volume[:, None]).sum(axis=0)
        imoment = (np.cross(positions, fE) * volume[:, None]).sum(axis=0)
        exp_loads = {"force_fluid": iforce, "moment_fluid": imoment,
                     "body_force": -iforce + vec["interior_momentum_rate"],
                     "body_moment": -imoment + vec["interior_angular_momentum_rate"]}
        if float(iv) <= 0 or any(not np.allclose(vec[n], v, rtol=ltol, atol=ltol) for n, v in exp_loads.items()):
            raise CaseConfigError("checkpoint closed-body loads disagree with its IB field or momentum rates")
        mode = str(np.asarray(ck["ib_force_mode"]).item())
        if mode != str(rt.config["ib"]["force_mode"]):
            raise CaseConfigError("checkpoint IB-force mode does not match the case")
        dmode = "accumulated_explicit" if mode == "accumulated_explicit" else "replacement"
        count = int(ck["ib_force_update_count"])
        exp_count = step if mode == "accumulated_explicit" else step * rt.inner_iterations
        if count != exp_count:
            raise CaseConfigError("checkpoint IB-force update count is inconsistent with its step")
    drv.time = t
    drv._step_count = step
    mtol = 1.0e-12
    if drv.motion is not None:
        em, ev = drv.motion(t)
        if not np.allclose(markers, em, rtol=0.0, atol=mtol):
            raise CaseConfigError("checkpoint markers disagree with prescribed motion at its time")
        if not np.allclose(mvel, ev, rtol=0.0, atol=mtol):
            raise CaseConfigError("checkpoint marker velocity disagrees with prescribed motion")
    else:
        if not np.allclose(markers, rt.reference_markers, rtol=0.0, atol=mtol):
            raise CaseConfigError("checkpoint markers disagree with the fixed body")
    drv.load_ib_state_dict({"force_mode": dmode, "force_E": fE, "force_L": fL, "force_update_count": count})
    drv.set_markers(markers, mvel)
    s = core.tensor(s_raw)
    s_old = core.tensor(so_raw)
    _, _, vc = drv.calc_vel(s)
    rec = rt.load_correction.evaluate(vc, sphere_center(rt.config, t))
    mt = 1.0e-11
    if not (np.allclose(rec.linear_momentum, vec["interior_momentum"], rtol=mt, atol=mt)
            and np.allclose(rec.angular_momentum, vec["interior_angular_momentum"], rtol=mt, atol=mt)
            and np.isclose(float(rec.indicator_volume), float(iv), rtol=mt, atol=mt)):
        raise CaseConfigError("checkpoint interior momentum disagrees with its flow field")
    rt.load_correction.restore(vec["interior_momentum"], vec["interior_angular_momentum"], iv)
    return step, s, s_old, vc, run_id


def _validated_history_prefix(path, start_step, rt):
    """ib_case.py:1227-1291."""
    if not path.exists():
        raise CaseConfigError("resume requires history.csv beside checkpoint.npz")
    kept, prev, saw = [], -1, False
    exp_updates = 1 if rt.config["ib"]["force_mode"] == "accumulated_explicit" else rt.inner_iterations
    with path.open("r", newline="", encoding="utf-8") as src:
        reader = csv.DictReader(src)
        if reader.fieldnames != HISTORY_COLUMNS:
            raise CaseConfigError("history.csv schema does not match this runner")
        for line, raw in enumerate(reader, start=2):
            try:
                if raw.get(None) or any(raw.get(n) is None for n in HISTORY_COLUMNS):
                    raise ValueError("wrong number of CSV fields")
                row = {}
                for n in HISTORY_COLUMNS:
                    v = float(raw[n])
                    if not np.isfinite(v):
                        raise ValueError(f"non-finite {n}")
                    if n in _INT_COLUMNS:
                        if int(v) != v:
                            raise ValueError(f"non-integer {n}")
                        row[n] = int(v)
                    else:
                        row[n] = v
                step = int(row["step"])
                if step <= prev:
                    raise ValueError("steps are not strictly increasing")
                if step > start_step:
                    if saw:
                        break
                    raise ValueError("history skipped the checkpoint step")
                et = step * rt.dt
                tol = max(1.0e-12, 16.0 * np.finfo(np.float64).eps * max(abs(et), rt.dt) * max(step, 1))
                if abs(float(row["time"]) - et) > tol:
                    raise ValueError("time is inconsistent with step and dt")
                if int(row["ib_force_updates"]) != exp_updates:
                    raise ValueError("IB update count is inconsistent with the configured method")
            except (KeyError, TypeError, ValueError) as exc:
                if saw:
                    break
                raise CaseConfigError(f"invalid history.csv row {line} before checkpoint: {exc}") from exc
            kept.append(row)
            prev = step
            if step == start_step:
                saw = True
    if not saw:
        raise CaseConfigError("history.csv does not contain the checkpoint step")
    return kept


def _atomic_history(path, rows):
    tmp = path.with_name(f".{path.name}.{uuid.uuid4().hex}.tmp")
    try:
        with tmp.open("x", newline="", encoding="utf-8") as h:
            w = csv.DictWriter(h, fieldnames=HISTORY_COLUMNS)
            w.writeheader()
            w.writerows(rows)
            h.flush()
            os.fsync(h.fileno())
        os.replace(tmp, path)
        _fsync_parent(path)
    finally:
        if tmp.exists():
            tmp.unlink()


def read_history(path):
    rows = []
    if not Path(path).exists():
        return rows
    with Path(path).open("r", newline="", encoding="utf-8") as h:
        for raw in csv.DictReader(h):
            rows.append({n: int(raw[n]) if n in _INT_COLUMNS else float(raw[n]) for n in HISTORY_COLUMNS})
    return rows


def _metadata(rt, run_id, outdir, start_step, resume):
    return {
        "result_schema": RESULT_SCHEMA_VERSION, "run_id": run_id,
        "created_utc": datetime.now(timezone.utc).isoformat(), "case_name": rt.config["name"],
        "case_kind": rt.config["case"]["kind"], "backend": BACKEND, "precision": _precision(rt),
        "python": platform.python_version(), "array_backend_version": torch.__version__,
        "config_fingerprint": config_fingerprint(rt.config), "layout_fingerprint": rt.layout_fingerprint,
        "config": rt.config, "mesh": rt.mesh_summary, "dt": rt.dt, "target_steps": rt.steps,
        "start_step": start_step, "resumed_from": str(Path(resume).resolve()) if resume else None,
        "output_directory": str(Path(outdir).resolve()), "source_sha256": rt.source_hashes,
        "closed_body_load": {
            "method": "Uhlmann interior-fluid momentum correction",
            "force": "-force_fluid + backward_difference(interior_momentum)",
            "moment": "-moment_fluid + backward_difference(interior_angular_momentum)",
            "indicator": "analytic sphere with one-cell linear smoothing", "moment_origin": [0.0, 0.0, 0.0]},
    }


def _read_resume_metadata(path, rt, run_id):
    try:
        md = json.loads(path.read_text(encoding="utf-8"))
    except (OSError, json.JSONDecodeError) as exc:
        raise CaseConfigError(f"cannot read existing metadata.json: {exc}") from exc
    exp = {"result_schema": RESULT_SCHEMA_VERSION, "run_id": run_id, "config_fingerprint": config_fingerprint(rt.config),
           "layout_fingerprint": rt.layout_fingerprint, "backend": BACKEND, "precision": _precision(rt)}
    for k, v in exp.items():
        if md.get(k) != v:
            raise CaseConfigError(f"existing metadata {k} does not match the checkpoint")
    if md.get("source_sha256") != rt.source_hashes:
        raise CaseConfigError("existing metadata source identity does not match this solver")
    if float(md.get("dt", -1.0)) != rt.dt:
        raise CaseConfigError("existing metadata dt does not match the checkpoint")
    return md


def _archive(outdir, paths):
    existing = [p for p in paths if p.exists()]
    if not existing:
        return None
    stamp = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%SZ")
    arch = outdir / "archive" / f"{stamp}_{uuid.uuid4().hex[:8]}"
    arch.mkdir(parents=True, exist_ok=False)
    for p in existing:
        os.replace(p, arch / p.name)
    _fsync_parent(arch)
    return arch


@dataclass
class RunResult:
    output_directory: Path
    start_step: int
    completed_step: int
    target_steps: int
    stopped: bool
    s: np.ndarray
    s_old: np.ndarray
    vc: np.ndarray
    ib_force_E: np.ndarray
    ib_force_L: np.ndarray
    body_force: np.ndarray
    body_moment: np.ndarray
    markers: np.ndarray
    marker_velocity: np.ndarray
    history: list


def run_case(config, output_directory, *, resume=None, overwrite=False, quiet=False, runtime=None):
    """ib_case.py:1413-1648 on the HIP device."""
    config = validate_config(config)
    if config["runtime"]["device"] != "hip":
        raise RuntimeError("this runner executes runtime.device='hip'; use the reference runner for cpu/cuda")
    outdir = Path(output_directory)
    outdir.mkdir(parents=True, exist_ok=True)
    hist_p, ck_p, md_p = outdir / "history.csv", outdir / "checkpoint.npz", outdir / "metadata.json"
    ff_p, snap_p = outdir / "final_field.npz", outdir / "snapshots"
    known = (hist_p, ck_p, md_p, ff_p, snap_p)
    rpath = Path(resume) if resume is not None else None
    if rpath is not None and rpath.resolve() != ck_p.resolve():
        raise CaseConfigError("resume must use checkpoint.npz in the requested output directory")
    if rpath is None and not overwrite and any(p.exists() for p in known):
        raise FileExistsError(f"{outdir} already contains run artifacts; pass overwrite=True or resume a checkpoint")
    t0 = time.perf_counter()
    rt = runtime if runtime is not None else build_runtime(config)
    if rpath is not None:
        start, s, s_old, vc, run_id = load_checkpoint(rpath, rt)
        if start >= rt.steps:
            raise CaseConfigError(f"target steps ({rt.steps}) must exceed checkpoint step ({start})")
        kept = _validated_history_prefix(hist_p, start, rt)
        prev_md = _read_resume_metadata(md_p, rt, run_id)
        _archive(outdir, (ff_p,))
        stale = tuple(p for p in snap_p.glob("field_*.npz") if p.stem[6:].isdigit() and int(p.stem[6:]) > start)
        _archive(outdir, stale)
    else:
        start = 0
        s, s_old, vc = initial_state(rt)
        run_id = str(uuid.uuid4())
        kept, prev_md = [], None
        if overwrite:
            _archive(outdir, known)
    md = _metadata(rt, run_id, outdir, start, rpath)
    if prev_md is not None:
        md["created_utc"] = prev_md["created_utc"]
        md["resume_count"] = int(prev_md.get("resume_count", 0)) + 1
        md["last_resume_utc"] = datetime.now(timezone.utc).isoformat()
    else:
        md["resume_count"] = 0
    md["setup_seconds"] = time.perf_counter() - t0
    _atomic_json(md_p, md)
    _atomic_history(hist_p, kept)
    handle = hist_p.open("a", newline="", encoding="utf-8")
    writer = csv.DictWriter(handle, fieldnames=HISTORY_COLUMNS)
    stop = {"requested": False}
    old = {}

    def request_stop(signum, _frame):
        stop["requested"] = True
        if not quiet:
            print(f"[signal] received {signum}; stopping after the current step", flush=True)

    for name in ("SIGTERM", "SIGINT", "SIGUSR1"):
        if hasattr(signal, name):
            num = int(getattr(signal, name))
            try:
                old[num] = signal.getsignal(num)
                signal.signal(num, request_stop)
            except (OSError, ValueError):
                pass
    done = start
    ti, si = time.perf_counter(), start
    drv = rt.driver
    if not quiet:
        print(f"CFDAgent-IB case={config['name']} backend={BACKEND} dtype={_precision(rt)} cells={rt.core.ncell:,} "
              f"edges={rt.core.nedge:,} markers={rt.reference_markers.shape[0]:,} steps={start + 1}..{rt.steps}",
              flush=True)
        print(f"output={outdir.resolve()}", flush=True)
    try:
        for step in range(start + 1, rt.steps + 1):
            drv.time = (step - 1) * rt.dt
            s, vc, s_old = drv.step(s, s_old, rt.dt, dt_old=rt.dt, inner=rt.inner_iterations)
            drv.time = step * rt.dt
            rt.load_correction.advance(vc, sphere_center(config, drv.time), rt.dt, drv.ib_force_E)
            done = step
            ce = int(config["output"]["checkpoint_every"])
            due = bool((ce and step % ce == 0) or step == rt.steps or stop["requested"])
            sample = step <= 3 or step % int(config["output"]["sample_every"]) == 0 or step == rt.steps or due
            if sample:
                torch.cuda.synchronize()
                now = time.perf_counter()
                ms = (now - ti) * 1000.0 / max(step - si, 1)
                ti, si = now, step
                row = diagnostics(rt, step, s, vc, ms)
                writer.writerow(row)
                handle.flush()
                if not quiet:
                    print(f"step={step:6d} t={row['time']:9.5f} Cd={row['Cd']:10.5f} slip={row['slip_max']:.4e} "
                          f"mass={row['mass_residual_inf']:.2e} CG={row['helmholtz_iterations']:4d} "
                          f"{row['milliseconds_per_step']:8.1f} ms/step", flush=True)
            if due:
                handle.flush()
                os.fsync(handle.fileno())
                save_checkpoint(ck_p, rt, run_id, step, s, s_old)
            fe = int(config["output"]["field_every"])
            if fe and step >= int(config["output"]["field_start_step"]) and step % fe == 0 and not (
                    step == rt.steps and config["output"]["save_field"]):
                snap_p.mkdir(exist_ok=True)
                save_field_snapshot(snap_p / f"field_{step:09d}.npz", rt, run_id, step, s, s_old, vc)
            if stop["requested"] and not due:
                if not sample:
                    torch.cuda.synchronize()
                    now = time.perf_counter()
                    ms = (now - ti) * 1000.0 / max(step - si, 1)
                    ti, si = now, step
                    writer.writerow(diagnostics(rt, step, s, vc, ms))
                handle.flush()
                os.fsync(handle.fileno())
                save_checkpoint(ck_p, rt, run_id, step, s, s_old)
            if stop["requested"]:
                break
    finally:
        handle.close()
        for num, h in old.items():
            try:
                signal.signal(num, h)
            except (OSError, ValueError):
                pass
    if done == rt.steps and bool(config["output"]["save_field"]):
        save_field_snapshot(ff_p, rt, run_id, done, s, s_old, vc)
    md.update({"completed_utc": datetime.now(timezone.utc).isoformat(), "completed_step": done,
               "complete": done == rt.steps})
    _atomic_json(md_p, md)
    load = rt.load_correction.latest_load
    if load is None:
        raise RuntimeError("closed-body load state is unavailable at run completion")
    if not quiet:
        print(f"CFDAgent-IB {'COMPLETE' if done == rt.steps else 'STOPPED'} step={done} time={drv.time:.6f}",
              flush=True)
    return RunResult(outdir, start, done, rt.steps, done != rt.steps, s.cpu().numpy(), s_old.cpu().numpy(),
                     vc[:rt.core.ncell].cpu().numpy(), drv.ib_force_E.cpu().numpy(), drv.ib_force_L.cpu().numpy(),
                     np.asarray(load.body_force), np.asarray(load.body_moment), drv.ib_markers.cpu().numpy(),
                     drv.ib_vel.cpu().numpy(), read_history(hist_p))
