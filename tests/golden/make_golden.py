"""Write golden vectors from the reference PyIB solver (run in the build container only).

The reference's ``make_euler_mesh.py`` imports h5py and meshio at module top
for its HDF5/VTU writers, which the solver path never calls; those two
modules are absent from this image, so empty placeholder modules are put in
``sys.modules`` before the import.  Nothing else of the reference is changed.

Fixtures written to tests/golden/ (NumPy .npz, loaded with allow_pickle=False):

* ``geometry_2level.npz`` -- IBCore geometry of a 2-level 2:1 nested mesh
  (Modify_geom interface corrections), ib_core_np.py:24-126.
* ``flow_mat_2level.npz`` -- flow_mat and CSR A@p of a seeded p (ib_core_np.py:168-175,
  ib_sparse.py:69-96).
* ``traj_<case>.npz`` -- short trajectories of the bundled example cases
  (ib_case.run_case): s, s_old, vc, forces after the run, plus every history row.

Usage:  python tests/golden/make_golden.py   (needs /root/reference)
"""
from __future__ import annotations

import json
import os
import sys
import tempfile
import types
from pathlib import Path

import numpy as np

REF = Path(os.environ.get("PYIB_REFERENCE", "/root/reference"))
OUT = Path(__file__).resolve().parent


def _import_reference():
    for name in ("h5py", "meshio"):
        sys.modules.setdefault(name, types.ModuleType(name))
    os.environ["IB_GPU"] = "0"
    os.environ["IB_FP32"] = "0"
    sys.path.insert(0, str(REF / "src" / "pyib" / "solver"))
    sys.path.insert(0, str(REF / "src" / "pyib"))
    import make_euler_mesh as M
    import ib_core_np
    import ib_sparse
    import ib_case
    return M, ib_core_np, ib_sparse, ib_case


def main():
    M, core_mod, sparse_mod, case_mod = _import_reference()
    # 2-level nested mesh: finest [-0.5,0.5]^3 at h=0.125 inside [-1,1]^3 at 0.25
    boxes = [M._build_box(("bounds", (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5))),
             M._build_box(("bounds", (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)))]
    mesh = M.make_nested_mesh(0.125, boxes, refinement_ratio=2)
    core = core_mod.IBCore(mesh)
    geo = {n: np.asarray(getattr(core, n)) for n in
           ("fpos", "cpos", "rf", "rf1", "rf2", "nf", "Af", "vol_ci", "AoL", "fe_sign_link")}
    np.savez_compressed(OUT / "geometry_2level.npz", spacing=np.float64(0.125),
                        inner=np.array([-0.5, 0.5]), outer=np.array([-1.0, 1.0]), **geo)
    rng = np.random.default_rng(20261203)
    p = rng.standard_normal(core.nedge)
    kw = dict(alpd=0.5, visc=0.01, rou=1.0, dt=0.01)
    w_free = core.flow_mat(p.copy(), **kw)
    A = sparse_mod.build_A(core, **kw, to_backend=False)
    np.savez_compressed(OUT / "flow_mat_2level.npz", p=p, w_flow_mat=w_free, w_csr=A @ p,
                        alpd=0.5, visc=0.01, rou=1.0, dt=0.01)
    cases = {
        "cpu_fixed_sphere_smoke": REF / "examples" / "cpu_fixed_sphere_smoke.json",
        "cpu_oscillating_sphere_smoke": REF / "examples" / "cpu_oscillating_sphere_smoke.json",
        "cpu_fixed_sphere_re100_d8_gate": REF / "examples" / "cpu_fixed_sphere_re100_d8_gate.json",
    }
    for name, path in cases.items():
        raw = json.loads(path.read_text())
        raw["output"]["save_field"] = False
        raw["output"]["sample_every"] = 1
        cfg = case_mod.validate_config(raw)
        with tempfile.TemporaryDirectory() as d:
            res = case_mod.run_case(cfg, d, quiet=True)
        hist = res.history
        cols = list(hist[0].keys())
        table = np.array([[row[c] for c in cols] for row in hist], dtype=np.float64)
        fields = dict(s=res.s, ib_force_L=res.ib_force_L, body_force=res.body_force, markers=res.markers)
        if res.s.shape[0] < 10000:       # the larger case keeps only s (vc, F_E follow from s, F_L)
            fields.update(s_old=res.s_old, vc=res.vc, ib_force_E=res.ib_force_E)
        np.savez_compressed(OUT / f"traj_{name}.npz", config=np.str_(json.dumps(cfg, sort_keys=True)),
                            history=table, columns=np.array(cols), **fields)
        print(name, "steps", len(hist), "CG", [int(r["helmholtz_iterations"]) for r in hist])


if __name__ == "__main__":
    main()
