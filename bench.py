"""Throughput of PyIB's time step on one MI355X: the D32 oscillating sphere.

Workload (the headline configuration): sphere of diameter D = 32 finest cells
oscillating along x, x(t) = A sin(2 pi f t), A = 0.1 D, Re0 = 0.2 (f = 3.6/2pi,
nu = 0.45), 3 nested 2:1 levels with the finest 192x160x160 (13,516,800 cells,
40,986,624 edge unknowns), 3217 Fibonacci markers moving every step, 400 steps
per period, 4 Crank-Nicolson inner iterations, Jacobi-PCG to the reference's
mixed tolerance (1e-10 / 1e-9, check cadence 20).  The run starts from rest at
t = 0; warm-up steps are run untimed, then steps are timed with the device
synchronised on both sides.

Prints one JSON line.  `value` is the steps/s of the plain float64 path (the
reference's arithmetic); the mixed-precision path (float32 PCG inside float64
reliable updates), its speedup and its body-force error against float64 over
the same steps are reported beside it.

  python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--precision both|mixed|float64]
                  [--budget-s 420] [--no-cpu-baseline] [--cpu-seconds 20] [--no-reference-case]

A float64 step at this size takes about 20 s, so the whole run is held to a wall
budget (`--budget-s`, counted from the start of the process, set-up included):
each precision runs at least one warm-up and one timed step, and at most the
requested numbers, as far as the budget allows.  The JSON line reports the steps
actually timed (`steps`, `warmup`) beside the requested ones.

`roofline.traffic` is the measured HBM traffic of one PCG iteration (rocprofv3 FETCH_SIZE
and WRITE_SIZE, scaled by a plain copy kernel), read from profiles/pcg_iteration_traffic_d32.md
when that summary is present.
"""
from __future__ import annotations

import time

T_START = time.perf_counter()

import argparse  # noqa: E402
import json  # noqa: E402
import math  # noqa: E402
import os  # noqa: E402
import platform  # noqa: E402
import sys  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-steps/s at D32 osc. sphere: FP64 vs FP32+FP64-refine HIP, at force err ≤1e-6"
HBM_PEAK_GBS = 8000.0
BYTES_PER_EDGE_ITER = {"float64": 104.0, "float32": 52.0}   # 13 words per edge per PCG iteration
# the reference's own CUDA float64 run of its D32 case (5-level concentric mesh): the first three
# history rows of checks/pressure-validation-20260925.zip runs/D32/history.csv
REFERENCE_D32_MS = {1: 11053.69720607996, 2: 18038.678898476064, 3: 16321.177490055561}
REFERENCE_D32_CG_MAX = {1: 2220, 2: 2220, 3: 2220}


def elapsed():
    return time.perf_counter() - T_START


def headline_config(steps=10, precision="float64", mixed_inner_tol=None):
    f = 3.6 / (2.0 * math.pi)
    T = 1.0 / f
    solver = {"sparse": False, "cg_check_interval": 20, "cg_max_iterations": 200000,
              "fail_on_cg_nonconvergence": True}
    if mixed_inner_tol is not None:
        solver["mixed_inner_tolerance"] = float(mixed_inner_tol)
    return {
        "schema_version": 1, "name": "bench_d32_oscillating_3level",
        "case": {"kind": "oscillating_sphere"},
        "runtime": {"device": "hip", "precision": precision},
        "mesh": {"kind": "nested_box", "cells_per_diameter": 32, "half_extents_d": [3.0, 2.5, 2.5], "levels": 3},
        "physics": {"reynolds": 0.2, "density": 1.0, "inflow_speed": 0.0},
        "time": {"dt": T / 400.0, "steps": steps, "inner_iterations": 4},
        "ib": {"radius": 0.25, "center": [0.0, 0.0, 0.0], "marker_spacing": 0.5 / 32,
               "force_mode": "accumulated_explicit", "implicit": False, "alpha": 1.0,
               "motion": {"kind": "oscillation", "axis": [1.0, 0.0, 0.0], "amplitude": 0.05, "frequency": f,
                          "phase": 0.0}},
        "solver": solver,
        "output": {"sample_every": 1, "checkpoint_every": 0, "save_field": False},
    }


class CGTimer:
    """HIP events (torch's current stream, where every kernel is launched) around each PCG block."""

    def __init__(self, torch, log=None, every_s=20.0):
        self.torch = torch
        self.pairs = []
        self.iters = 0
        self.log = log
        self.every_s = every_s
        self._last = time.perf_counter()

    def begin(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def end(self, start, n):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.pairs.append((start, e, n))
        self.iters += int(n[0])
        now = time.perf_counter()
        if self.log is not None and now - self._last > self.every_s:
            self._last = now
            self.log(f"[bench]   ... {self.iters} PCG iterations so far ({n[1]})")

    def reset(self):
        self.pairs, self.iters = [], 0

    def summary(self):
        self.torch.cuda.synchronize()
        by = {}
        for a, b, (n, prec) in self.pairs:
            t, k = by.get(prec, (0.0, 0))
            by[prec] = (t + a.elapsed_time(b), k + n)
        return {p: {"ms": t, "iters": k, "ms_per_iter": t / max(k, 1)} for p, (t, k) in by.items()}


def build_mesh(log, config=None):
    import pyib_hip  # noqa: F401
    from pyib_hip import case as HC

    t0 = time.perf_counter()
    mesh, dx, summary = HC.build_mesh(HC.validate_config(config or headline_config(1)))
    log(f"[bench] mesh {summary['total_cells']:,} cells in {time.perf_counter() - t0:.1f} s "
        f"(process at {elapsed():.0f} s)")
    return mesh


def run_gpu(precision, steps, warmup, log, mesh, deadline, mixed_inner_tol=None, config=None, keep_host=True):
    """Warm-up and timed steps of one precision, as many as the requested numbers and the wall
    deadline (perf_counter seconds) allow; at least one of each."""
    import torch
    from pyib_hip import case as HC

    cfg = HC.validate_config(config or headline_config(steps + warmup, precision, mixed_inner_tol))
    if config is not None:
        cfg["runtime"]["precision"] = precision
    t1 = time.perf_counter()
    rt = HC.build_runtime(cfg, mesh=mesh)
    log(f"[bench] runtime ({precision}) {rt.core.nedge:,} edges, {rt.reference_markers.shape[0]} markers "
        f"in {time.perf_counter() - t1:.1f} s (process at {elapsed():.0f} s)")
    drv = rt.driver
    timer = CGTimer(torch, log)
    drv.cg_timer = timer
    s, s_old, vc = HC.initial_state(rt)
    iters, iters_max, forces, step_s = [], [], [], []

    def one(step):
        nonlocal s, s_old, vc
        t = time.perf_counter()
        drv.time = (step - 1) * rt.dt
        s, vc, s_old = drv.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        drv.time = step * rt.dt
        rt.load_correction.advance(vc, HC.sphere_center(cfg, drv.time), rt.dt, drv.ib_force_E)
        iters.append(drv.helmholtz_iterations_sum_last_step)
        iters_max.append(drv.helmholtz_iterations_last_step)
        forces.append(float(rt.load_correction.latest_load.body_force[0]))
        torch.cuda.synchronize()
        step_s.append(time.perf_counter() - t)

    step = 0
    while True:
        step += 1
        one(step)
        log(f"[bench] {precision} warm-up step {step}: CG iterations {iters[-1]} in {step_s[-1]:.1f} s "
            f"(process at {elapsed():.0f} s)")
        if step >= warmup or time.perf_counter() + 2.2 * max(step_s) > deadline:
            break
    n_warm = step
    torch.cuda.synchronize()
    timer.reset()
    t = time.perf_counter()
    n_timed = 0
    while True:
        step += 1
        one(step)
        n_timed += 1
        log(f"[bench] {precision} step {step}: CG iterations {iters[-1]} at {time.perf_counter() - t:.1f} s "
            f"(process at {elapsed():.0f} s)")
        if n_timed >= steps or time.perf_counter() + 1.15 * max(step_s[n_warm:]) > deadline:
            break
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    cg = timer.summary()
    out = {"steps_per_s": n_timed / wall, "ms_per_step": 1000.0 * wall / n_timed, "steps": n_timed,
           "warmup": n_warm, "cg_iters_per_step": sum(iters[n_warm:]) / n_timed,
           "cg_iters_max_inner": iters_max, "cg_iters_sum": iters, "step_ms": [1000.0 * x for x in step_s],
           "cg": cg, "nedge": rt.core.nedge, "ncell": rt.core.ncell, "nface": rt.core.nface,
           "body_force_x": forces, "markers": int(rt.reference_markers.shape[0])}
    if keep_host:
        out.update(geometry=rt.core.geo, config=cfg)
    del rt, drv
    torch.cuda.empty_cache()
    return out


def measured_traffic(prec):
    """HBM bytes of one PCG iteration from the committed rocprofv3 summary, or None."""
    import re
    path = os.path.join(ROOT, "profiles", "pcg_iteration_traffic_d32.md")
    if not os.path.exists(path):
        return None
    m = re.search(rf"traffic_bytes_per_iteration {prec} = ([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_facts():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"nproc": affinity, "cpu_count": os.cpu_count(), "cpu_model": model,
            "OPENBLAS_NUM_THREADS": os.environ.get("OPENBLAS_NUM_THREADS", "unset"),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS", "unset")}


def cpu_baseline(mesh, geo, cfg, cg_iters_per_step, seconds, log):
    """The oracle (NumPy, one core for bincount/ufuncs) on a bounded sample of the same workload:
    as many PCG iterations of the matrix-free operator (ib_driver_np.py:386-403) as fit in
    ``seconds`` (at least two), plus one inner iteration's work outside the CG (calc_vel,
    conv_diff, right-hand side), scaled to the GPU run's PCG iterations per step."""
    import numpy as np
    from oracle.pyib_ref import case as OC

    t0 = time.perf_counter()
    ort = OC.build_runtime(cfg, mesh=mesh, sparse=False, geometry=geo)
    log(f"[bench] oracle runtime in {time.perf_counter() - t0:.1f} s (process at {elapsed():.0f} s)")
    drv, core = ort.driver, ort.core
    s, s_old, vc = OC.initial_state(ort)
    dt = ort.dt
    rng = np.random.default_rng(0)
    res = rng.standard_normal(core.nedge)
    res[drv.dir_mask] = 0.0
    x, r = s.copy(), res.copy()
    z = r * drv.prec
    p = z.copy()
    eta = float(z.dot(r))
    n_it = 0
    t = time.perf_counter()
    while True:
        w = drv.apply_A(p, dt)
        w[drv.dir_mask] = 0.0
        pw = float(p.dot(w))
        alf = eta / pw if pw != 0.0 else 0.0
        x += alf * p
        r -= alf * w
        z = r * drv.prec
        eta_new = float(z.dot(r))
        p = z + (eta_new / eta if eta != 0.0 else 0.0) * p
        eta = eta_new
        n_it += 1
        if n_it >= 2 and time.perf_counter() - t > seconds:
            break
    t_iter = (time.perf_counter() - t) / n_it
    t = time.perf_counter()
    _, U, vc = drv.calc_vel(s)
    drv.conv_diff(U, vc)
    core.rot(core.integr_c2f(vc[:core.ncell]))
    t_outside = time.perf_counter() - t
    # per step: 4 inner iterations, each a calc_vel + conv_diff + right-hand side + the true-residual
    # product (one more A application), plus the predictor's calc_vel + conv_diff
    t_step = cg_iters_per_step * t_iter + 4 * (t_outside + t_iter) + 2 * t_outside
    log(f"[bench] oracle: {n_it} PCG iterations at {t_iter:.2f} s each, {t_outside:.2f} s per inner "
        f"iteration outside CG (process at {elapsed():.0f} s)")
    sample = (f"oracle (NumPy restatement of the reference; matrix-free flow_mat because the CSR operator "
              f"of solver.sparse would need ~62 GB) on the same mesh: {n_it} PCG iterations (A.p + vector "
              f"updates) and one calc_vel + conv_diff + rhs, scaled to {cg_iters_per_step:.0f} PCG iterations "
              "per step (the GPU float64 run's count) plus 6 calc_vel/conv_diff and 4 true-residual products "
              "per step")
    return dict({"value": 1.0 / t_step, "unit": "time-steps/s", "cores": 1, "kind": "port", "sample": sample,
                 "pcg_iterations_timed": n_it, "s_per_pcg_iteration": t_iter, "s_per_step": t_step,
                 "threads_note": "NumPy bincount, ufuncs and fancy indexing run on one thread; only dot "
                                 "products may use OpenBLAS threads"}, **host_facts())


def run_reference_case(log, deadline, steps=3):
    """The reference's own D32 case file (5-level concentric mesh, 3,637,824 edges, A = 0.0111 D,
    560 steps per period) on the HIP float64 path from rest, against its CUDA float64 run."""
    import copy
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", "reference_runs.json")
    if not os.path.exists(path):
        return None
    ref = json.load(open(path))["D32"]["case"]
    cfg = copy.deepcopy(ref)
    cfg["runtime"]["device"] = "hip"
    cfg["time"]["steps"] = steps
    cfg["solver"]["sparse"] = False
    mesh = build_mesh(log, cfg)
    r = run_gpu("float64", steps - 1, 1, log, mesh, deadline, config=cfg, keep_host=False)
    done = r["warmup"] + r["steps"]
    ours = r["step_ms"][:done]
    theirs = [REFERENCE_D32_MS[k] for k in range(1, done + 1)]
    cg_ok = all(r["cg_iters_max_inner"][k - 1] == REFERENCE_D32_CG_MAX[k] for k in range(1, done + 1))
    return {"case": "reference D32 case file (sphere_concentric, 5 levels, finest 64^3, 1,179,648 cells, "
                    "3,637,824 edges, 3217 markers), steps 1..%d from rest" % done,
            "hip_float64_ms_per_step": ours, "reference_cuda_float64_ms_per_step": theirs,
            "speedup_vs_reference_cuda": float(np.sum(theirs) / np.sum(ours)),
            "cg_iterations_max_inner": r["cg_iters_max_inner"][:done],
            "cg_iterations_match_reference": cg_ok,
            "note": "reference timings are the CuPy/CUDA float64 run's milliseconds_per_step (GPU model not "
                    "recorded); ours include the same diagnostics-free step only"}


def kernels_only(n_iter, precision, log):
    """N PCG iterations of the headline operator on a seeded right-hand side, then a plain
    copy kernel of 2**26 doubles (read 512 MiB, write 512 MiB) to calibrate HBM counters."""
    import ctypes as C
    import numpy as np
    import torch
    import pyib_hip  # noqa: F401
    from pyib_hip import _native as N
    from pyib_hip import case as HC
    from pyib_hip.core import dev_stream, ptr

    cfg = HC.validate_config(headline_config(1, precision))
    mesh, _, _ = HC.build_mesh(cfg)
    rt = HC.build_runtime(cfg, mesh=mesh)
    drv, g = rt.driver, rt.driver.grid
    log(f"[kernels] {g.nedge:,} internal edges ({rt.core.nedge:,} mesh edges)")
    rng = np.random.default_rng(0)
    p0 = rng.standard_normal(rt.core.nedge)
    p0[drv.dir_mask] = 0.0
    # a right-hand side in the operator's range (A is singular off the Dirichlet edges)
    res_d = rt.core.flow_mat(rt.core.tensor(p0), alpd=drv.ad, visc=drv.visc, rou=drv.rou, dt=rt.dt, mask_dir=True)
    prec = "float32" if precision == "mixed" else "float64"
    m, ws = g.struct(prec), g.work_struct(prec)
    cg = drv._cgi32 if prec == "float32" else drv._cgi
    dt = torch.float32 if prec == "float32" else torch.float64
    r_i = g.to_internal(res_d, dtype=dt)
    z_i = torch.zeros(g.nedge, dtype=dt, device="cuda")
    N.call("pyib_grid_cg_begin", m, ws, C.byref(cg), ptr(r_i), ptr(z_i), 0.0, 1e-30, dev_stream())
    N.call("pyib_grid_cg_iterate", m, ws, C.byref(cg), float(drv.rou), float(drv.visc), 0.5 * rt.dt, int(n_iter), 0,
           dev_stream())
    torch.cuda.synchronize()
    a = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    N.call("pyib_convert", N.F64, ptr(a), N.F64, ptr(b), C.c_int64(a.numel()), dev_stream())
    torch.cuda.synchronize()
    log(f"[kernels] ran {n_iter} {prec} PCG iterations and one 2^26-double copy")
    print(json.dumps({"kernels_only": n_iter, "precision": prec, "nedge_internal": g.nedge,
                      "nedge": rt.core.nedge, "ncell_internal": g.ncell}), flush=True)
    return 0


def period_harmonics(d, steps_per_period, log):
    """Mixed precision against float64 over one full oscillation period of the oscillating sphere
    at cells_per_diameter d (3 nested levels, finest 8D x 6D x 6D, Re0 = 0.2, A = 0.1 D, dt =
    T/steps_per_period), from rest: the least-squares harmonic vector (a, b, c) of body_force_x,
    a cos(wt + phase) + b sin(wt + phase) + c over the period (the reference's validation fit,
    tools/check_pressure_validation.py:32-38), for each precision, and their relative difference."""
    import numpy as np
    import torch
    from pyib_hip import case as HC

    f = 3.6 / (2.0 * math.pi)
    T = 1.0 / f
    out = {}
    mesh = None
    for prec in ("float64", "mixed"):
        cfg = headline_config(steps_per_period, prec)
        cfg["name"] = f"period_d{d}"
        cfg["mesh"] = {"kind": "nested_box", "cells_per_diameter": d, "half_extents_d": [4.0, 3.0, 3.0], "levels": 3}
        cfg["ib"]["marker_spacing"] = 0.5 / d
        cfg["time"]["dt"] = T / steps_per_period
        cfg = HC.validate_config(cfg)
        if mesh is None:
            mesh, _, _ = HC.build_mesh(cfg)
        rt = HC.build_runtime(cfg, mesh=mesh)
        drv = rt.driver
        s, s_old, vc = HC.initial_state(rt)
        times, fx, iters = [], [], 0
        t0 = time.perf_counter()
        for step in range(1, steps_per_period + 1):
            drv.time = (step - 1) * rt.dt
            s, vc, s_old = drv.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
            drv.time = step * rt.dt
            load = rt.load_correction.advance(vc, HC.sphere_center(cfg, drv.time), rt.dt, drv.ib_force_E)
            times.append(drv.time)
            fx.append(float(load.body_force[0]))
            iters += drv.helmholtz_iterations_sum_last_step
            if step % 20 == 0:
                log(f"[period] {prec} step {step}/{steps_per_period}: Fx {fx[-1]:.12e}, "
                    f"{time.perf_counter() - t0:.0f} s")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        omega, phase = 2.0 * math.pi * f, float(cfg["ib"]["motion"]["phase"])
        ang = omega * np.asarray(times) + phase
        A = np.column_stack([np.cos(ang), np.sin(ang), np.ones_like(ang)])
        coeff = np.linalg.lstsq(A, np.asarray(fx), rcond=None)[0]
        out[prec] = {"harmonic_abc": [float(c) for c in coeff], "wall_s": wall, "cg_iterations": int(iters),
                     "body_force_x": fx}
        out["mesh"] = {"cells": int(rt.core.ncell), "edges": int(rt.core.nedge),
                       "markers": int(rt.reference_markers.shape[0])}
        del rt, drv, s, s_old, vc
        torch.cuda.empty_cache()
    h64, hmx = np.asarray(out["float64"]["harmonic_abc"]), np.asarray(out["mixed"]["harmonic_abc"])
    f64, fmx = np.asarray(out["float64"]["body_force_x"]), np.asarray(out["mixed"]["body_force_x"])
    res = {"case": f"oscillating sphere D{d}, 3 nested levels, finest {8 * d}x{6 * d}x{6 * d}, one period "
                   f"from rest in {steps_per_period} steps", "mesh": out["mesh"],
           "float64": {k: v for k, v in out["float64"].items() if k != "body_force_x"},
           "mixed": {k: v for k, v in out["mixed"].items() if k != "body_force_x"},
           "harmonic_relerr": float(np.linalg.norm(hmx - h64) / np.linalg.norm(h64)),
           "force_history_max_relerr": float(np.max(np.abs(fmx - f64)) / np.max(np.abs(f64))),
           "mixed_speedup": out["float64"]["wall_s"] / out["mixed"]["wall_s"]}
    print(json.dumps(res), flush=True)
    return 0


def _kernel_label(name):
    """`pyib::k_name<T>` (or the name itself) from a demangled kernel signature."""
    import re
    m = re.search(r"(k_[A-Za-z0-9_]+(?:<[^()]*?>)?)\(", name)
    return f"pyib::{m.group(1)}" if m and "pyib::" in name else name.split("(")[0].replace("void ", "")


def _rocpd_rows(path, sql):
    import sqlite3
    con = sqlite3.connect(f"file:{path}?mode=ro", uri=True)
    try:
        return con.execute(sql).fetchall()
    finally:
        con.close()


def summarize_profile(trace_db, fetch_db, write_db, n_iter, precision, nedge, command):
    """Markdown summary of one PCG iteration from the rocprofv3 databases of ``--kernels-only``
    runs: kernel averages (kernel trace), HBM bytes per dispatch (FETCH_SIZE / WRITE_SIZE, in
    KiB, scaled by what they read for the plain copy kernel of 2^26 doubles that the run ends
    with), and the algorithmic fraction of the 8 TB/s roofline."""
    stats = {}
    for name, calls, total in _rocpd_rows(trace_db, "select name, count(*), sum(duration) from kernels group by name"):
        lab = _kernel_label(name)
        c, t = stats.get(lab, (0, 0.0))
        stats[lab] = (c + int(calls), t + float(total) / 1e3)           # ns -> us

    def counter(db):
        out = {}
        for name, val in _rocpd_rows(db, "select kernel_name, value from counters_collection"):
            lab = _kernel_label(name)
            s, k = out.get(lab, (0.0, 0))
            out[lab] = (s + float(val), k + 1)
        return {k: s / n for k, (s, n) in out.items()}                    # KiB per dispatch

    fetch, write = counter(fetch_db), counter(write_db)
    copy = "pyib::k_convert<double, double>"
    fs = 2.0 ** 19 / fetch[copy] if fetch.get(copy) else 1.0               # 512 MiB read
    ws = 2.0 ** 19 / write[copy] if write.get(copy) else 1.0               # 512 MiB written
    iter_k = sorted(((lab, c, t) for lab, (c, t) in stats.items() if c >= n_iter and lab.startswith("pyib::")),
                    key=lambda x: -x[2])
    per_iter = sum(t for _, _, t in iter_k) / n_iter
    bpe = BYTES_PER_EDGE_ITER["float32" if precision == "mixed" else "float64"]
    rows, rd_tot, wr_tot = [], 0.0, 0.0
    for lab, c, t in iter_k:
        avg = t / c
        share = (t / n_iter) / per_iter
        rd = fetch.get(lab, 0.0) * 1024 * fs * (c / n_iter)
        wr = write.get(lab, 0.0) * 1024 * ws * (c / n_iter)
        rd_tot, wr_tot = rd_tot + rd, wr_tot + wr
        gbs = (rd + wr) / ((t / n_iter) * 1e-6) / 1e9 if t else 0.0
        rows.append(f"| `{lab}` | {c} | {avg:.1f} | {100 * share:.1f}% | {rd / 1e6:.0f} | {wr / 1e6:.0f} | "
                    f"{gbs:.0f} | {gbs / HBM_PEAK_GBS:.2f} |")
    alg = bpe * nedge
    prec_word = "float32" if precision == "mixed" else "float64"
    lines = [
        f"# One {prec_word} PCG iteration of the D32 headline operator (kernel trace and HBM traffic)", "",
        "Workload: the D32 oscillating-sphere case, 3 nested levels, finest 192x160x160: "
        f"{nedge:,} edge unknowns in the structured layout. `bench.py --kernels-only N` runs N PCG iterations of "
        "its Helmholtz operator on a right-hand side in the operator's range, then one plain copy kernel of "
        "2^26 doubles, which calibrates the HBM counters.", "",
        "Commands (one MI355X, after `cd /tmp && export TMPDIR=/tmp PYIB_HIP_SETUP_CACHE=/tmp/pyib_cache`; "
        "`$R` is the repository root):", "", *[f"    {c}" for c in command], "",
        f"Counter calibration: the copy kernel reads and writes 512 MiB each; FETCH_SIZE is scaled by {fs:.3f} and "
        f"WRITE_SIZE by {ws:.3f} below.", "",
        "| kernel | calls | avg us | share of iteration | measured HBM read MB | write MB | measured GB/s | of 8 TB/s |",
        "|---|---:|---:|---:|---:|---:|---:|---:|", *rows, "",
        f"Per iteration: {per_iter:.1f} us (sum of the kernel times over {n_iter} iterations); measured HBM traffic "
        f"{(rd_tot + wr_tot) / 1e9:.3f} GB ({(rd_tot + wr_tot) / alg:.2f}x the algorithmic bytes), i.e. "
        f"{(rd_tot + wr_tot) / (per_iter * 1e-6) / 1e12:.2f} TB/s; algorithmic bytes {alg / 1e9:.3f} GB "
        f"({bpe:.0f} B per edge unknown), i.e. {alg / (per_iter * 1e-6) / 1e12:.2f} TB/s = "
        f"{alg / (per_iter * 1e-6) / 1e9 / HBM_PEAK_GBS:.3f} of the 8 TB/s roofline.", "",
        f"traffic_bytes_per_iteration {prec_word} = {rd_tot + wr_tot:.6e}", "",
        "Other kernels in the same trace (set-up of the right-hand side and the copy kernel):", "",
        *[f"- `{lab}`: {c} call(s), {t / c:.1f} us avg" for lab, (c, t) in
          sorted(stats.items(), key=lambda kv: -kv[1][1]) if (lab, c, t) not in iter_k],
    ]
    return "\n".join(lines) + "\n"


def roofline(r, prec):
    cg = r["cg"].get(prec) or next(iter(r["cg"].values()))
    bytes_iter = BYTES_PER_EDGE_ITER[prec] * r["nedge"]
    achieved = bytes_iter / (cg["ms_per_iter"] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(prec),
            "kernel": f"one {prec} PCG iteration (G1 and V passes with p.Ap, G3 pass fused with "
                      "r -= alpha A p and r.prec.r, direction fused with x += alpha p)",
            "ms_per_iteration": cg["ms_per_iter"], "pcg_iterations_timed": cg["iters"],
            "algorithmic_bytes_per_iteration": bytes_iter,
            "algorithmic_bytes_per_edge": BYTES_PER_EDGE_ITER[prec]}


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", choices=("mixed", "float64", "both"), default="both")
    p.add_argument("--budget-s", type=float, default=420.0,
                   help="wall seconds for the whole run, counted from process start")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--no-reference-case", action="store_true",
                   help="skip the reference's own D32 case (steps 1-3 against its CUDA timings)")
    p.add_argument("--kernels-only", type=int, default=0, help="profiling: N PCG iterations + a copy kernel")
    p.add_argument("--summarize-profile", nargs=3, metavar=("TRACE_DB", "FETCH_DB", "WRITE_DB"),
                   help="print the markdown summary of --kernels-only rocprofv3 runs (kernel trace, FETCH_SIZE, "
                        "WRITE_SIZE databases); --kernels-only gives their iteration count, --precision theirs")
    p.add_argument("--profile-command", action="append", default=[],
                   help="a command line to quote in the summary (repeatable)")
    p.add_argument("--period-harmonics", type=int, default=0, metavar="D",
                   help="mixed against float64 over one oscillation period at cells_per_diameter D "
                        "(finest 8Dx6Dx6D, 3 levels)")
    p.add_argument("--period-steps", type=int, default=200, help="steps per period of --period-harmonics")
    p.add_argument("--mixed-inner-tol", type=float, default=None,
                   help="float32 residual reduction between float64 reliable updates (default: the schema's)")
    a = p.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: the solver has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        return 2
    if a.steps < 1 or a.warmup < 1:
        print("bench.py needs --steps >= 1 and --warmup >= 1 (the first step from rest is never timed)",
              file=sys.stderr)
        return 2
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    if a.summarize_profile:
        n_it = a.kernels_only or 100
        print(summarize_profile(*a.summarize_profile, n_it, a.precision, 40986624, a.profile_command), end="")
        return 0
    if a.period_harmonics:
        return period_harmonics(a.period_harmonics, a.period_steps, log)
    if a.kernels_only:
        return kernels_only(a.kernels_only, "mixed" if a.precision == "mixed" else "float64", log)
    precs = ["float64", "mixed"] if a.precision == "both" else [a.precision]
    deadline_all = T_START + a.budget_s
    cpu_reserve = 0.0 if a.no_cpu_baseline else a.cpu_seconds + 45.0
    ref_reserve = 0.0 if a.no_reference_case else 40.0
    results = {}
    mesh = build_mesh(log)
    for i, prec in enumerate(precs):
        left = deadline_all - time.perf_counter() - cpu_reserve - ref_reserve
        share = left * (0.62 if (prec == "float64" and len(precs) == 2) else 1.0)
        results[prec] = run_gpu(prec, a.steps, a.warmup, log, mesh, time.perf_counter() + share,
                                a.mixed_inner_tol)
    head = "float64" if "float64" in results else "mixed"
    r = results[head]
    inner_prec = "float64" if head == "float64" else "float32"
    line = {
        "metric": METRIC, "value": r["steps_per_s"], "unit": "time-steps/s", "n_gpus": 1, "steps": r["steps"],
        "warmup": r["warmup"], "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if head == "float64" else "f32+f64",
        "data": "synthetic: the case itself from rest (no dataset), prescribed sphere oscillation",
        "config": {"workload": "D32 oscillating sphere, 3-level nested box, finest 192x160x160 "
                               f"({r['ncell']:,} cells, {r['nedge']:,} edges), {r['markers']} markers, Re0=0.2, "
                               "A=0.1D, dt=T/400, 4 inner CN iterations, PCG tol 1e-10/1e-9",
                   "precision": head, "cg_iters_per_step": r["cg_iters_per_step"]},
        "steps_requested": a.steps, "warmup_requested": a.warmup, "budget_s": a.budget_s,
        "roofline": roofline(r, inner_prec),
    }
    for prec, rr in results.items():
        line[f"{prec}_steps_per_s"] = rr["steps_per_s"]
        line[f"{prec}_ms_per_step"] = rr["ms_per_step"]
        line[f"{prec}_steps_timed"] = rr["steps"]
        line[f"{prec}_cg_iters_per_step"] = rr["cg_iters_per_step"]
    if "mixed" in results:
        line["roofline_mixed"] = roofline(results["mixed"], "float32")
    if "mixed" in results and "float64" in results:
        import numpy as np
        fm, f64 = results["mixed"]["body_force_x"], results["float64"]["body_force_x"]
        n = min(len(fm), len(f64))
        den = max(abs(x) for x in f64[:n])
        line["mixed_vs_float64_force_relerr"] = float(np.max(np.abs(np.array(fm[:n]) - np.array(f64[:n]))) / den)
        line["mixed_vs_float64_steps_compared"] = n
        line["mixed_speedup"] = results["mixed"]["steps_per_s"] / results["float64"]["steps_per_s"]
    if not a.no_reference_case and deadline_all - time.perf_counter() > cpu_reserve + 25.0:
        try:
            line["reference_case_d32"] = run_reference_case(log, deadline_all - cpu_reserve)
        except Exception as exc:  # the headline line is still printed
            line["reference_case_d32"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not a.no_cpu_baseline:
        left = deadline_all - time.perf_counter() - 30.0
        line["cpu_baseline"] = cpu_baseline(mesh, r["geometry"], r["config"], r["cg_iters_per_step"],
                                            max(5.0, min(a.cpu_seconds, left)), log)
        line["float64_vs_cpu_baseline"] = r["steps_per_s"] / line["cpu_baseline"]["value"]
    line["wall_s"] = elapsed()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
