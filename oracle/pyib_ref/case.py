"""Oracle restatement of the case setup, step loop and per-step diagnostics.

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_case.py``:
``fibonacci_sphere`` 554-563, ``sphere_center`` 587-603, ``_layout_fingerprint``
490-510, ``config_fingerprint`` 455-467, ``build_runtime`` 633-809,
``_initial_state`` 812-820, ``diagnostics`` 1129-1224 and the step loop of
``run_case`` 1517-1543.  The harmonic fit is ``tools/check_pressure_validation.py:32-38``.
The config is taken already normalised (the product's ``validate_config``);
mesh kinds: uniform_box, sphere_wake, sphere_concentric and the additive
nested_box.
"""
from __future__ import annotations

import copy
import hashlib
import json
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import mesh as M
from .core import Core
from .driver import Driver
from .immersed import ImmersedBoundary
from .interior import ClosedSphereLoadCorrection
from .kinematics import Oscillation

LAYOUT_NAMES = ("nodes", "cells", "boundary_faces", "boundary_fam", "pos_ref", "dx_uniform",
                "index_cell", "faces_to_nodes", "faces_to_cells", "edges_to_nodes", "faces_to_edges",
                "bc_node", "bc_edge", "bc_face", "bc_cell", "interface_pairs", "link_fe_insert")


def fibonacci_sphere(count, radius, centre):
    """ib_case.py:554-563."""
    i = np.arange(count, dtype=float)
    phi = np.arccos(1.0 - 2.0 * (i + 0.5) / count)
    theta = np.pi * (1.0 + np.sqrt(5.0)) * i
    unit = np.stack([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)], axis=1)
    return np.asarray(centre, dtype=float) + radius * unit


def layout_fingerprint(mesh, markers, widths):
    """ib_case.py:490-510."""
    d = hashlib.sha256()
    for name in LAYOUT_NAMES:
        a = np.ascontiguousarray(np.asarray(getattr(mesh, name)))
        d.update(name.encode() + b"\0")
        d.update(a.dtype.str.encode("ascii") + b"\0")
        d.update(json.dumps(a.shape).encode("ascii") + b"\0")
        d.update(a.tobytes(order="C"))
    for name, v in (("markers", markers), ("marker_widths", widths)):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        d.update(name.encode() + b"\0")
        d.update(json.dumps(a.shape).encode("ascii") + b"\0")
        d.update(a.tobytes(order="C"))
    return d.hexdigest()


def config_fingerprint(config):
    """ib_case.py:455-467."""
    c = copy.deepcopy(config)
    c.pop("name", None)
    c.pop("output", None)
    c["time"].pop("steps", None)
    return hashlib.sha256(json.dumps(c, sort_keys=True, separators=(",", ":"), allow_nan=False)
                          .encode("utf-8")).hexdigest()


def rotating_sphere_velocity(markers, centre, axis, spin_ratio, radius, reference_speed):
    """ib_case.py:566-584: rigid surface velocity omega x (x - c), |omega| = spin_ratio U / R."""
    pts = np.asarray(markers, float)
    d = np.asarray(axis, float)
    n = float(np.linalg.norm(d))
    if pts.ndim != 2 or pts.shape[1] != 3 or d.shape != (3,) or not np.isfinite(n) or n == 0.0:
        raise ValueError("markers must be (N,3) and the rotation axis a finite non-zero three-vector")
    omega = d / n * float(spin_ratio) * float(reference_speed) / float(radius)
    return np.cross(omega, pts - np.asarray(centre, float))


def sphere_center(config, t):
    """ib_case.py:587-603."""
    c = np.asarray(config["ib"]["center"], float)
    mo = config["ib"]["motion"]
    if mo["kind"] != "oscillation":
        return c
    ang = 2.0 * np.pi * float(mo["frequency"]) * float(t) + float(mo["phase"])
    return c + float(mo["amplitude"]) * np.sin(ang) * np.asarray(mo["axis"], float)


def harmonic_fit(times, values, omega, phase):
    """tools/check_pressure_validation.py:32-38: least squares a cos + b sin + c."""
    ang = omega * np.asarray(times) + phase
    A = np.column_stack([np.cos(ang), np.sin(ang), np.ones_like(ang)])
    coeff, _, rank, _ = np.linalg.lstsq(A, np.asarray(values), rcond=None)
    if rank != 3:
        raise ValueError("harmonic fit is rank deficient")
    return coeff


def mesh_for(config):
    """ib_case.py:640-688 mesh selection (+ nested_box)."""
    mc = config["mesh"]
    radius = float(config["ib"]["radius"])
    kind = mc["kind"]
    if kind == "uniform_box":
        dx = float(mc["dx"])
        return M.make_nested_mesh(dx, M.uniform_box(mc["bounds"])), dx
    ddx = int(mc["cells_per_diameter"])
    dx = 2.0 * radius / ddx
    if kind == "sphere_wake":
        return M.make_nested_mesh(M.DIAMETER / ddx, M.wake_boxes()), dx
    if kind == "sphere_concentric":
        return M.make_nested_mesh(M.DIAMETER / ddx, M.concentric_boxes()), dx
    if kind == "nested_box":
        return M.make_nested_mesh(M.DIAMETER / ddx, M.nested_box_boxes(mc["half_extents_d"], int(mc["levels"]))), dx
    raise ValueError(kind)


@dataclass
class Runtime:
    config: dict
    mesh: Any
    core: Core
    driver: Driver
    ib: ImmersedBoundary
    load: ClosedSphereLoadCorrection
    markers: np.ndarray
    widths: np.ndarray
    reference_force: float
    reference_speed: float
    dx: float
    dt: float
    steps: int
    inner: int


def uses_cavity(config):
    """ib_case.py:717-718 (sphere_concentric), extended to a quiescent nested_box."""
    kind = config["mesh"]["kind"]
    return kind == "sphere_concentric" or (kind == "nested_box" and float(config["physics"]["inflow_speed"]) == 0.0)


def build_runtime(config, mesh=None, sparse=None, geometry=None):
    """ib_case.py:633-809 (explicit IB only)."""
    ibc = config["ib"]
    radius = float(ibc["radius"])
    centre = np.asarray(ibc["center"], float)
    if mesh is None:
        mesh, dx = mesh_for(config)
    else:
        dx = 2.0 * radius / int(config["mesh"]["cells_per_diameter"]) if config["mesh"]["kind"] != "uniform_box" \
            else float(config["mesh"]["dx"])
    core = Core(mesh, geometry=geometry)
    ph = config["physics"]
    U = float(ph["inflow_speed"])
    Re = float(ph["reynolds"])
    rho = float(ph["density"])
    mo = ibc["motion"]
    if U != 0.0:
        nu = abs(U) * 2.0 * radius / Re
        uref = abs(U)
    else:
        uref = 2.0 * np.pi * float(mo["frequency"]) * float(mo["amplitude"])
        nu = uref * 2.0 * radius / Re
    sv = config["solver"]
    drv = Driver(core, mesh, visc=rho * nu, rou=rho, cg_max_iterations=int(sv["cg_max_iterations"]),
                 fail_on_cg_nonconvergence=bool(sv["fail_on_cg_nonconvergence"]),
                 cg_check_interval=int(sv["cg_check_interval"]))
    if uses_cavity(config):
        drv.set_cavity(U_lid=0.0, spanwise_zerograd=False)
    dt = float(config["time"]["dt"])
    if (sv["sparse"] if sparse is None else sparse):
        drv.enable_sparse(dt)
    count = max(12, int(round(4.0 * np.pi * radius ** 2 / float(ibc["marker_spacing"]) ** 2)))
    markers = fibonacci_sphere(count, radius, centre)
    widths = np.full(count, (4.0 * np.pi * radius ** 2 * dx / count) ** (1.0 / 3.0))
    ib = ImmersedBoundary(mesh)
    mode = "accumulated_explicit" if ibc["force_mode"] == "accumulated_explicit" else "replacement"
    if mo["kind"] == "oscillation":
        motion = Oscillation(markers, axis=mo["axis"], A=float(mo["amplitude"]), freq=float(mo["frequency"]),
                             phase=float(mo["phase"]))
        m0, v0 = motion(0.0)
        drv.set_ib(ib, m0, widths, v0, force_mode=mode)
        drv.set_motion(motion, t0=0.0)
    elif mo["kind"] == "rotation":
        drv.set_ib(ib, markers, widths, rotating_sphere_velocity(markers, centre, mo["axis"], float(mo["spin_ratio"]),
                                                                 radius, uref), force_mode=mode)
    else:
        drv.set_ib(ib, markers, widths, np.zeros((count, 3)), force_mode=mode)
    drv.ib_alpha = float(ibc["alpha"])
    drv.ib_implicit = bool(ibc["implicit"])                        # ib_case.py:766
    mt = float(mo["amplitude"]) * np.abs(np.asarray(mo["axis"], float)) if mo["kind"] == "oscillation" else np.zeros(3)
    load = ClosedSphereLoadCorrection(core.cpos[:core.ncell], 1.0 / core.vol_ci, reference_center=centre,
                                      maximum_translation=mt, radius=radius, smoothing_width=dx, density=rho,
                                      force_support_width=1.5 * ib.hh)
    fref = 0.5 * rho * uref ** 2 * (np.pi * radius ** 2)
    return Runtime(config, mesh, core, drv, ib, load, markers, widths, fref, uref, dx, dt,
                   int(config["time"]["steps"]), int(config["time"]["inner_iterations"]))


def initial_state(rt):
    """ib_case.py:812-820."""
    core = rt.core
    s = float(rt.config["physics"]["inflow_speed"]) * core.epos[:, 1] * core.te[:, 2]
    s, _, vc = rt.driver.calc_vel(s)
    rt.load.initialize(vc[:core.ncell], sphere_center(rt.config, rt.driver.time))
    return s, s.copy(), vc


def diagnostics(rt, step, s, vc):
    """ib_case.py:1129-1224 (without wall-clock), returns the history row."""
    core, drv, load = rt.core, rt.driver, rt.load.latest_load
    bf = load.body_force
    mv = rt.ib.interp(vc[:core.ncell], drv.ib_markers)
    slip = np.linalg.norm(mv - drv.ib_vel, axis=1)
    speed = np.linalg.norm(vc[:core.ncell], axis=1)
    ff = drv.last_face_flux
    if ff is None:
        _, ff, _ = drv.calc_vel(s)
    div = core.vol_ci * core.cdiv(ff)
    lag = (drv.ib_force_L * (drv.ib_ds ** 3)[:, None]).sum(axis=0)
    den = max(float(np.linalg.norm(lag)), float(np.linalg.norm(load.force_fluid)), 1e-300)
    bal = float(np.linalg.norm(load.force_fluid - lag)) / den
    co = bf / rt.reference_force
    row = {"step": int(step), "time": float(drv.time), "Cd": float(co[0]), "Cy": float(co[1]), "Cz": float(co[2])}
    for name, v in (("body_force", bf), ("force_fluid", load.force_fluid), ("body_moment", load.body_moment),
                    ("moment_fluid", load.moment_fluid), ("interior_momentum", load.interior_momenta.linear_momentum),
                    ("interior_angular_momentum", load.interior_momenta.angular_momentum),
                    ("interior_momentum_rate", load.linear_momentum_rate),
                    ("interior_angular_momentum_rate", load.angular_momentum_rate)):
        for k, ax in enumerate("xyz"):
            row[f"{name}_{ax}"] = float(v[k])
    sm = float(speed.max())
    row.update({
        "interior_indicator_volume": float(load.interior_momenta.indicator_volume),
        "slip_max": float(slip.max()), "slip_rms": float(np.sqrt(np.mean(slip ** 2))),
        "speed_max": sm, "cfl": sm * rt.dt / rt.dx, "mass_residual_inf": float(np.max(np.abs(div))),
        "force_balance_relerr": bal, "ib_force_updates": int(drv.ib_force_updates_last_step),
        "helmholtz_iterations": int(drv.helmholtz_iterations_last_step),
        "helmholtz_relative_residual": float(drv.helmholtz_relative_residual_last_step),
        "helmholtz_tolerance_ratio": float(drv.helmholtz_tolerance_ratio_last_step),
    })
    return row


def run(rt, steps=None, record=None):
    """ib_case.py:1517-1531 step loop; returns (s, s_old, vc, rows) with a row every step."""
    s, s_old, vc = initial_state(rt)
    rows = []
    n = rt.steps if steps is None else int(steps)
    for step in range(1, n + 1):
        rt.driver.time = (step - 1) * rt.dt
        s, vc, s_old = rt.driver.step(s, s_old, rt.dt, dt_old=rt.dt, inner=rt.inner)
        rt.driver.time = step * rt.dt
        rt.load.advance(vc[:rt.core.ncell], sphere_center(rt.config, rt.driver.time), rt.dt, rt.driver.ib_force_E)
        row = diagnostics(rt, step, s, vc)
        row["helmholtz_iterations_sum"] = int(rt.driver.helmholtz_iterations_sum_last_step)
        rows.append(row)
        if record is not None:
            record(step, s, vc)
    return s, s_old, vc, rows
