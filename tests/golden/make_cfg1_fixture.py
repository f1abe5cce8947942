"""Write the oracle's run of the smallest full configuration (cfg1) as a fixture.

cfg1: fixed sphere, Re 20, D = 8 finest cells, 2 nested levels, finest 64x48x48 (8D x 6D x 6D),
276,480 cells, 859,248 edges, 201 markers, dt = 0.5 h / U, 10 steps, float64, CG 1e-10 / 1e-9.

The oracle (oracle/pyib_ref, the NumPy restatement of the reference) runs it twice, as the
reference can: with the matrix-free flow_mat (solver.sparse false, the path the HIP solver is
compared with) and with the assembled CSR matrix (solver.sparse true).  Their difference is the
floor below which the two reference paths do not agree with each other.

Writes tests/golden/cfg1_oracle.npz (allow_pickle=False):
  config                       the validated case (JSON text)
  columns, history             the 41 history columns of every step, matrix-free run
  history_csr                  the same rows of the CSR run
  stride, s_sub, omega_sub     s and the edge vorticity omega = rot(integr_c2f(vc)) after step 10
                               on every stride-th edge, matrix-free run
  s_sub_csr, omega_sub_csr     the same entries of the CSR run
  s_norm, omega_norm           full-field L2 norms, matrix-free run
  floor_s, floor_omega         full-field relative L2 of CSR vs matrix-free

The two runs take about an hour each on one core; they run in two processes.
Usage:  python tests/golden/make_cfg1_fixture.py
"""
from __future__ import annotations

import json
import multiprocessing as mp
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent / "cfg1_oracle.npz"
STRIDE = 32
STEPS = 10


def _config():
    sys.path.insert(0, str(ROOT))
    import pyib_hip  # noqa: F401
    from tests.helpers import full_size_config
    return full_size_config("cfg1", steps=STEPS, device="cpu")


def _run(sparse):
    sys.path.insert(0, str(ROOT))
    from oracle.pyib_ref import case as OC
    cfg = _config()
    t0 = time.perf_counter()
    rt = OC.build_runtime(cfg, sparse=sparse)
    tag = "csr" if sparse else "matrix-free"
    say = lambda step, s, vc: print(f"[cfg1 {tag}] step {step} at {time.perf_counter() - t0:.0f} s", flush=True)
    s, _, vc, rows = OC.run(rt, STEPS, record=say)
    omega = rt.core.rot(rt.core.integr_c2f(vc[:rt.core.ncell]))
    return rows, np.asarray(s), np.asarray(omega)


def main():
    with np.load(Path(__file__).resolve().parent / "traj_cpu_fixed_sphere_smoke.npz", allow_pickle=False) as g:
        cols = [str(c) for c in g["columns"]]       # the reference's history.csv columns
    with mp.get_context("spawn").Pool(2) as pool:
        (rows, s, om), (rows_c, s_c, om_c) = pool.map(_run, [False, True])
    hist = np.array([[float(r.get(c, 0.0)) for c in cols] for r in rows])
    hist_c = np.array([[float(r.get(c, 0.0)) for c in cols] for r in rows_c])
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    np.savez_compressed(
        OUT, config=json.dumps(_config(), sort_keys=True), columns=np.array(cols), history=hist,
        history_csr=hist_c, stride=np.int64(STRIDE), s_sub=s[::STRIDE], omega_sub=om[::STRIDE],
        s_sub_csr=s_c[::STRIDE], omega_sub_csr=om_c[::STRIDE], s_norm=np.float64(np.linalg.norm(s)),
        omega_norm=np.float64(np.linalg.norm(om)), floor_s=np.float64(rel(s_c, s)),
        floor_omega=np.float64(rel(om_c, om)))
    print(f"wrote {OUT}: floor s {rel(s_c, s):.3e}, omega {rel(om_c, om):.3e}")


if __name__ == "__main__":
    main()
