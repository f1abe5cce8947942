"""Throughput of PyIB's time step on one MI355X: the D32 oscillating sphere.

Workload (the headline configuration): sphere of diameter D = 32 finest cells
oscillating along x, x(t) = A sin(2 pi f t), A = 0.1 D, Re0 = 0.2 (f = 3.6/2pi,
nu = 0.45), 3 nested 2:1 levels with the finest 192x160x160 (13,516,800 cells,
40,986,624 edge unknowns), 3217 Fibonacci markers moving every step, 400 steps
per period, 4 Crank-Nicolson inner iterations, Jacobi-PCG to the reference's
mixed tolerance (1e-10 / 1e-9, check cadence 20).  The run starts from rest at
t = 0; W warm-up steps are run untimed, then K steps are timed with the device
synchronised on both sides.

Prints one JSON line.  `value` is the steps/s of the mixed-precision path
(float32 PCG inside float64 reliable updates) when it is run (the default runs
both), else of the float64 path; both paths' steps/s, the mixed-vs-float64
force error and the speedup are reported beside it.

  python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--precision both|mixed|float64]
                  [--no-cpu-baseline] [--cpu-seconds 20]

`roofline.traffic` is the measured HBM traffic of one PCG iteration (rocprofv3 FETCH_SIZE
and WRITE_SIZE, scaled by a plain copy kernel), read from profiles/pcg_iteration_traffic_d32.md
when that summary is present.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-steps/s at D32 osc. sphere: FP64 vs FP32+FP64-refine HIP, at force err ≤1e-6"
HBM_PEAK_GBS = 8000.0
BYTES_PER_EDGE_ITER = {"float64": 104.0, "float32": 52.0}   # 13 words per edge per PCG iteration


def headline_config(steps=10, precision="float64", mixed_inner_tol=None):
    f = 3.6 / (2.0 * math.pi)
    T = 1.0 / f
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
        "solver": {"sparse": False, "cg_check_interval": 20, "cg_max_iterations": 200000,
                   "fail_on_cg_nonconvergence": True},
        "output": {"sample_every": 1, "checkpoint_every": 0, "save_field": False},
    } | ({} if mixed_inner_tol is None else {"solver": {"sparse": False, "cg_check_interval": 20,
                                                      "cg_max_iterations": 200000,
                                                      "fail_on_cg_nonconvergence": True,
                                                      "mixed_inner_tolerance": float(mixed_inner_tol)}})


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


def build_mesh(log):
    import pyib_hip  # noqa: F401
    from pyib_hip import case as HC

    t0 = time.perf_counter()
    mesh, dx, summary = HC.build_mesh(HC.validate_config(headline_config(1)))
    log(f"[bench] mesh {summary['total_cells']:,} cells in {time.perf_counter() - t0:.1f} s")
    return mesh


def run_gpu(precision, steps, warmup, log, mesh, mixed_inner_tol=None):
    import torch
    from pyib_hip import case as HC

    cfg = HC.validate_config(headline_config(steps + warmup, precision, mixed_inner_tol))
    t1 = time.perf_counter()
    rt = HC.build_runtime(cfg, mesh=mesh)
    log(f"[bench] runtime ({precision}) {rt.core.nedge:,} edges, {rt.reference_markers.shape[0]} markers "
        f"in {time.perf_counter() - t1:.1f} s")
    drv = rt.driver
    timer = CGTimer(torch, log)
    drv.cg_timer = timer
    s, s_old, vc = HC.initial_state(rt)
    iters, forces = [], []

    def one(step):
        nonlocal s, s_old, vc
        drv.time = (step - 1) * rt.dt
        s, vc, s_old = drv.step(s, s_old, rt.dt, rt.dt, inner=rt.inner_iterations)
        drv.time = step * rt.dt
        rt.load_correction.advance(vc, HC.sphere_center(cfg, drv.time), rt.dt, drv.ib_force_E)
        iters.append(drv.helmholtz_iterations_sum_last_step)
        forces.append(float(rt.load_correction.latest_load.body_force[0]))

    for step in range(1, warmup + 1):
        tw = time.perf_counter()
        one(step)
        log(f"[bench] warmup step {step}: CG iterations {iters[-1]} in {time.perf_counter() - tw:.1f} s")
    torch.cuda.synchronize()
    timer.reset()
    t = time.perf_counter()
    for step in range(warmup + 1, warmup + steps + 1):
        one(step)
        log(f"[bench] step {step}: CG iterations {iters[-1]} at {time.perf_counter() - t:.1f} s")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    cg = timer.summary()
    out = {"steps_per_s": steps / elapsed, "ms_per_step": 1000.0 * elapsed / steps,
           "cg_iters_per_step": sum(iters[warmup:]) / steps, "cg": cg, "nedge": rt.core.nedge,
           "ncell": rt.core.ncell, "nface": rt.core.nface, "body_force_x": forces, "markers":
           int(rt.reference_markers.shape[0]), "mesh": mesh, "geometry": rt.core.geo, "config": cfg}
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


def cpu_baseline(mesh, geo, cfg, cg_iters_per_step, seconds, log):
    """Oracle (NumPy, one core for bincount/ufuncs) on a bounded sample of the same workload."""
    import numpy as np
    from oracle.pyib_ref import case as OC

    t0 = time.perf_counter()
    ort = OC.build_runtime(cfg, mesh=mesh, sparse=False, geometry=geo)
    log(f"[bench] oracle runtime in {time.perf_counter() - t0:.1f} s")
    drv, core = ort.driver, ort.core
    s, s_old, vc = OC.initial_state(ort)
    dt = ort.dt
    rng = np.random.default_rng(0)
    res = rng.standard_normal(core.nedge)
    res[drv.dir_mask] = 0.0
    # (1) one PCG iteration = one matrix-free A.p (ib_driver_np.py:387) + the vector updates (389-403)
    p = res * drv.prec
    t = time.perf_counter()
    w = drv.apply_A(p, dt)
    t_A = time.perf_counter() - t
    x, r = s.copy(), res.copy()
    t = time.perf_counter()
    w[drv.dir_mask] = 0.0
    alf = float(r.dot(p * drv.prec)) / float(p.dot(w))
    x += alf * p
    r -= alf * w
    z = r * drv.prec
    eta = z.dot(r)
    p = z + (eta / max(eta, 1e-300)) * p
    t_vec = time.perf_counter() - t
    t_iter = t_A + t_vec
    # (2) per-inner-iteration work outside CG: calc_vel, conv_diff, rhs, true-residual product
    t = time.perf_counter()
    _, U, vc = drv.calc_vel(s)
    conv, diff = drv.conv_diff(U, vc)
    core.rot(core.integr_c2f(vc[:core.ncell]))
    t_inner = time.perf_counter() - t + t_A
    t_step = cg_iters_per_step * t_iter + 4 * t_inner + 2 * (t_inner - t_A)
    log(f"[bench] oracle: {t_iter:.2f} s per PCG iteration, {t_inner:.2f} s per inner iteration outside CG")
    sample = ("oracle (NumPy/SciPy restatement; matrix-free flow_mat because the CSR operator of solver.sparse "
              "would need ~62 GB) on the same mesh: 1 PCG iteration (A.p + vector updates) and 1 calc_vel + "
              f"conv_diff + rhs, scaled to {cg_iters_per_step:.0f} PCG iterations per step (the GPU run's count) "
              "and 6 calc_vel/conv_diff per step")
    return {"value": 1.0 / t_step, "unit": "time-steps/s", "cores": 1, "kind": "port", "sample": sample,
            "s_per_pcg_iteration": t_iter, "s_per_step": t_step}


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


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", choices=("mixed", "float64", "both"), default="both")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--kernels-only", type=int, default=0, help="profiling: N PCG iterations + a copy kernel")
    p.add_argument("--mixed-inner-tol", type=float, default=None,
                   help="float32 residual reduction between float64 reliable updates (default: the schema's)")
    a = p.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        return 2
    log = lambda msg: print(msg, file=sys.stderr, flush=True)
    if a.kernels_only:
        return kernels_only(a.kernels_only, "mixed" if a.precision == "mixed" else "float64", log)
    precs = ["float64", "mixed"] if a.precision == "both" else [a.precision]
    results = {}
    mesh = build_mesh(log)
    for prec in precs:
        results[prec] = run_gpu(prec, a.steps, a.warmup, log, mesh, a.mixed_inner_tol)
    head = "mixed" if "mixed" in results else "float64"
    r = results[head]
    inner_prec = "float32" if head == "mixed" else "float64"
    cg = r["cg"].get(inner_prec) or next(iter(r["cg"].values()))
    bytes_iter = BYTES_PER_EDGE_ITER[inner_prec] * r["nedge"]
    achieved = bytes_iter / (cg["ms_per_iter"] * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": r["steps_per_s"], "unit": "time-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic: the case itself from rest (no dataset), prescribed sphere oscillation",
        "config": {"workload": "D32 oscillating sphere, 3-level nested box, finest 192x160x160 "
                               f"({r['ncell']:,} cells, {r['nedge']:,} edges), {r['markers']} markers, Re0=0.2, "
                               "A=0.1D, dt=T/400, 4 inner CN iterations, PCG tol 1e-10/1e-9",
                   "precision": head, "cg_iters_per_step": r["cg_iters_per_step"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(inner_prec),
                     "kernel": f"one {inner_prec} PCG iteration (G1 and V passes with p.Ap, G3 pass fused with "
                               "r -= alpha A p and r.prec.r, direction fused with x += alpha p)",
                     "ms_per_iteration": cg["ms_per_iter"],
                     "algorithmic_bytes_per_iteration": bytes_iter},
    }
    for prec, rr in results.items():
        line[f"{prec}_steps_per_s"] = rr["steps_per_s"]
        line[f"{prec}_cg_iters_per_step"] = rr["cg_iters_per_step"]
    if "mixed" in results and "float64" in results:
        fm, f64 = results["mixed"]["body_force_x"], results["float64"]["body_force_x"]
        import numpy as np
        den = max(abs(x) for x in f64)
        line["mixed_vs_float64_force_relerr"] = float(np.max(np.abs(np.array(fm) - np.array(f64))) / den)
        line["mixed_speedup"] = results["mixed"]["steps_per_s"] / results["float64"]["steps_per_s"]
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(r["mesh"], r["geometry"], r["config"], r["cg_iters_per_step"],
                                            a.cpu_seconds, log)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
