"""Shared test helpers: configs of the parity cases and array comparisons."""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CAPTURE_MANAGER = None      # pytest's capture manager, set by conftest.pytest_configure


def progress(msg):
    """Print a progress line of a long test (a step done) past pytest's output capture."""
    if CAPTURE_MANAGER is not None:
        with CAPTURE_MANAGER.global_and_fixture_disabled():
            print(msg, flush=True)
    else:
        print(msg, flush=True)


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


def parity_tol(floor, base=1e-12, factor=3.0):
    """Tolerance of a float64 parity check: the reference's own CSR (solver.sparse) and
    matrix-free flow_mat paths differ by `floor` on the same case; a HIP result must stay
    within max(base, factor*floor) of the matrix-free oracle."""
    return max(base, factor * floor)


def parity_report(tag, err, floor, base=1e-12, factor=3.0):
    """Print a parity measurement with its ratio to the oracle's own two-path floor and
    return whether it is within parity_tol."""
    ratio = err / floor if floor > 0 else float("inf")
    progress(f"[gpu] {tag}: {err:.3e} (floor {floor:.3e}, ratio {ratio:.2f}, "
             f"allowance {parity_tol(floor, base, factor):.3e})")
    return err <= parity_tol(floor, base, factor)


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


def full_size_config(which, steps=2, precision="float64", device="hip"):
    """The full-size configurations as PyIB JSON (nested_box meshes, radius 0.25 = D/2).

    cfg1: fixed sphere, Re 20, D8, 2 levels, finest 64x48x48 (8D x 6D x 6D), dt = 0.5 h/U, 10 steps.
    cfg2: fixed sphere, Re 100, D16, 3 levels, finest 160x128x128 (10D x 8D x 8D), dt = 0.5 h/U.
    cfg3: oscillating sphere, Re0 0.2, A = 0.1D, D32, 3 levels, finest 192x160x160, dt = T/400.
    cfg4: oscillating sphere, Re0 0.2, A = 0.1D, D40, 3 levels, finest 240x200x200, dt = T/400.
    cfg5: oscillating sphere, Re0 0.2, A = 0.1D, D16, 3 levels, finest 128x96x96, dt = T/200.
    """
    import math
    import pyib_hip.case as HC
    f = 3.6 / (2.0 * math.pi)
    osc = {"kind": "oscillation", "axis": [1.0, 0.0, 0.0], "amplitude": 0.05, "frequency": f, "phase": 0.0}
    levels = 3
    if which == "cfg1":
        d, half, ph, ib, dt = 8, [4.0, 3.0, 3.0], {"reynolds": 20.0, "inflow_speed": 1.0}, {}, 0.5 * (0.5 / 8)
        case, levels = {"kind": "fixed_sphere"}, 2
    elif which == "cfg3":
        d, half, ph, ib, dt = 32, [3.0, 2.5, 2.5], {"reynolds": 0.2, "inflow_speed": 0.0}, {"motion": osc}, (1 / f) / 400
        case = {"kind": "oscillating_sphere"}
    elif which == "cfg2":
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
    raw = {"name": f"full_{which}", "case": case, "runtime": {"device": device, "precision": precision},
           "mesh": {"kind": "nested_box", "cells_per_diameter": d, "half_extents_d": half, "levels": levels},
           "physics": ph, "time": {"dt": dt, "steps": steps, "inner_iterations": 4},
           "ib": dict({"marker_spacing": 0.5 / d, "force_mode": "accumulated_explicit"}, **ib),
           "solver": {"sparse": False, "cg_check_interval": 20, "cg_max_iterations": 200000,
                      "fail_on_cg_nonconvergence": True}}
    return HC.validate_config(raw)


VECTOR_GROUPS = ("body_force", "force_fluid", "body_moment", "moment_fluid", "interior_momentum",
                 "interior_angular_momentum", "interior_momentum_rate", "interior_angular_momentum_rate")
EXACT_COLUMNS = ("step", "ib_force_updates", "helmholtz_iterations")


MOMENT_OF = {"body_moment": "body_force", "moment_fluid": "force_fluid",
             "interior_angular_momentum": "interior_momentum",
             "interior_angular_momentum_rate": "interior_momentum_rate"}


def compare_history(rows, hist, cols, *, rtol=1e-10, resid_rtol=1e-3, radius=0.25,
                    skip=("milliseconds_per_step",)):
    """Every history column of ``rows`` (dicts, or a (history, columns) pair) against the reference
    rows ``hist`` with columns ``cols``.  Rules per column:

    * step, ib_force_updates, helmholtz_iterations: identical; time: equal to 1e-15 relative;
    * the (x, y, z) columns of each load vector and Cd/Cy/Cz: |a - b| <= rtol * the largest
      magnitude of the group (the cross-stream components of a symmetric case are round-off);
      a moment group's magnitude is at least ``radius`` times that of its force group (the
      moments of a symmetric case are round-off too);
    * interior_indicator_volume, slip_max, slip_rms, speed_max, cfl: rtol relative;
    * mass_residual_inf (max |vol_ci div U|, zero in exact arithmetic, of order speed / dx when
      wrong): |a - b| <= 1e-10 times the reference's speed_max;
      force_balance_relerr: both <= 1e-9 (the reference's run-time invariant);
    * helmholtz_relative_residual, helmholtz_tolerance_ratio (depend on where each CG's round-off
      leaves the residual): resid_rtol relative.

    Returns {column: worst deviation / its allowance}; raises AssertionError naming every
    column over its allowance."""
    if isinstance(rows, tuple):
        h2, c2 = rows
        get = lambda c: np.asarray(h2[:, list(c2).index(c)], float)
    else:
        get = lambda c: np.array([float(r[c]) for r in rows])
    ref = lambda c: np.asarray(hist[:, list(cols).index(c)], float)
    report, bad = {}, []

    def chk(c, dev, allow):
        ratio = float(np.max(dev / allow)) if np.size(dev) else 0.0
        report[c] = ratio
        if not ratio <= 1.0:
            bad.append(f"{c}: {ratio:.3g} x allowance")

    groups = {g: [f"{g}_{a}" for a in "xyz"] for g in VECTOR_GROUPS}
    groups["C"] = ["Cd", "Cy", "Cz"]
    grouped = {c: g for g, cs in groups.items() for c in cs}

    def group_scale(g):
        sc = max(np.abs(ref(x)).max() for x in groups[g])
        if g in MOMENT_OF:      # a moment is also measured against its force times the radius
            sc = max(sc, radius * max(np.abs(ref(x)).max() for x in groups[MOMENT_OF[g]]))
        return max(sc, 1e-300)
    speed = ref("speed_max")
    for c in cols:
        if c in skip:
            continue
        a, b = get(c), ref(c)
        if a.shape != b.shape:
            bad.append(f"{c}: {a.shape[0]} rows vs {b.shape[0]}")
            continue
        if c in EXACT_COLUMNS:
            chk(c, np.abs(a - b), 0.5)
        elif c == "time":
            chk(c, np.abs(a - b), 1e-15 * np.abs(b).max() + 1e-300)
        elif c in grouped:
            chk(c, np.abs(a - b), rtol * group_scale(grouped[c]))
        elif c == "mass_residual_inf":
            chk(c, np.abs(a - b), 1e-10 * speed + 1e-300)
        elif c == "force_balance_relerr":
            chk(c, np.maximum(a, b), 1e-9)
        elif c in ("helmholtz_relative_residual", "helmholtz_tolerance_ratio"):
            chk(c, np.abs(a - b), resid_rtol * np.abs(b) + 1e-300)
        else:
            chk(c, np.abs(a - b), rtol * np.abs(b).max() + 1e-300)
    assert not bad, "history columns over their allowance: " + "; ".join(bad)
    return report
