"""Shared test helpers: configs of the parity cases and array comparisons."""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def golden_config(name):
    with golden(f"traj_{name}.npz") as g:
        return json.loads(str(g["config"]))


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def small_case(kind="oscillating", device="hip", precision="float64", steps=3, mesh=None, **over):
    """Small parity cases (normalised by the product's validate_config)."""
    raw = {"name": f"parity_{kind}", "runtime": {"device": device, "precision": precision},
           "time": {"dt": 0.01, "steps": steps, "inner_iterations": 4},
           "solver": {"sparse": False, "cg_max_iterations": 2000}}
    if kind == "oscillating":
        raw["case"] = {"kind": "oscillating_sphere"}
        raw["ib"] = {"marker_spacing": 0.125, "motion": {"kind": "oscillation", "axis": [1, 0, 0],
                                                          "amplitude": 0.05, "frequency": 0.5729577951308232}}
        raw["physics"] = {"reynolds": 0.2, "inflow_speed": 0.0}
        raw["mesh"] = mesh or {"kind": "nested_box", "cells_per_diameter": 8, "half_extents_d": [1.5, 1.25, 1.25],
                               "levels": 2}
    else:
        raw["case"] = {"kind": "fixed_sphere"}
        raw["ib"] = {"marker_spacing": 0.0625}
        raw["physics"] = {"reynolds": 100.0, "inflow_speed": 1.0}
        raw["mesh"] = mesh or {"kind": "nested_box", "cells_per_diameter": 8, "half_extents_d": [1.5, 1.25, 1.25],
                               "levels": 2}
    for k, v in over.items():
        raw.setdefault(k, {}).update(v)
    import pyib_hip.case as HC
    return HC.validate_config(raw)


def parity_tol(floor, base=1e-12, factor=10.0):
    """Tolerance of a float64 parity check: the reference's own CSR (solver.sparse) and
    matrix-free flow_mat paths differ by `floor` on the same case; a HIP result must stay
    within max(base, factor*floor) of the matrix-free oracle."""
    return max(base, factor * floor)


FORCE_COLUMNS = ("body_force_x", "body_force_y", "body_force_z")


def force_history(rows, cols=FORCE_COLUMNS):
    """(steps, 3) body-force history from history rows (dicts) or a (history, columns) pair."""
    if isinstance(rows, tuple):
        hist, names = rows
        return np.stack([hist[:, list(names).index(c)] for c in cols], axis=1)
    return np.array([[r[c] for c in cols] for r in rows])


def force_rel(ours, ref):
    """Relative L2 of a force history: ||F - F_ref|| / ||F_ref|| over every step and component (the
    cross-stream components of a symmetric case are near zero, so each component is not scaled by itself)."""
    return rel(np.asarray(ours), np.asarray(ref))


def full_size_config(which, steps=2, precision="float64"):
    """The full-size configurations as PyIB JSON (nested_box meshes, radius 0.25 = D/2).

    cfg2: fixed sphere, Re 100, D16, 3 levels, finest 160x128x128 (10D x 8D x 8D), dt = 0.5 h/U.
    cfg4: oscillating sphere, Re0 0.2, A = 0.1D, D40, 3 levels, finest 240x200x200, dt = T/400.
    cfg5: oscillating sphere, Re0 0.2, A = 0.1D, D16, 3 levels, finest 128x96x96, dt = T/200.
    """
    import math
    import pyib_hip.case as HC
    f = 3.6 / (2.0 * math.pi)
    osc = {"kind": "oscillation", "axis": [1.0, 0.0, 0.0], "amplitude": 0.05, "frequency": f, "phase": 0.0}
    if which == "cfg2":
        d, half, ph, ib, dt = 16, [5.0, 4.0, 4.0], {"reynolds": 100.0, "inflow_speed": 1.0}, {}, 0.5 * (0.5 / 16)
        case = {"kind": "fixed_sphere"}
    elif which == "cfg4":
        d, half, ph, ib, dt = 40, [3.0, 2.5, 2.5], {"reynolds": 0.2, "inflow_speed": 0.0}, {"motion": osc}, (1 / f) / 400
        case = {"kind": "oscillating_sphere"}
    elif which == "cfg5":
        d, half, ph, ib, dt = 16, [4.0, 3.0, 3.0], {"reynolds": 0.2, "inflow_speed": 0.0}, {"motion": osc}, (1 / f) / 200
        case = {"kind": "oscillating_sphere"}
    else:
        raise ValueError(which)
    raw = {"name": f"full_{which}", "case": case, "runtime": {"device": "hip", "precision": precision},
           "mesh": {"kind": "nested_box", "cells_per_diameter": d, "half_extents_d": half, "levels": 3},
           "physics": ph, "time": {"dt": dt, "steps": steps, "inner_iterations": 4},
           "ib": dict({"marker_spacing": 0.5 / d, "force_mode": "accumulated_explicit"}, **ib),
           "solver": {"sparse": False, "cg_check_interval": 20, "cg_max_iterations": 200000,
                      "fail_on_cg_nonconvergence": True}}
    return HC.validate_config(raw)
