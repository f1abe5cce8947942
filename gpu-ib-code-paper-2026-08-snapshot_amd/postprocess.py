"""Offline pressure recovery of a saved field on the GPU.

Counterpart of ``recover_pressure`` of the reference's offline post-processor
(``src/pyib/solver/ib_postprocess.py:49-105``), which reconstructs the whole-field pressure from
the final inner iteration's momentum residual (its gradient is ``momentum_residual / dt``,
``ib_postprocess.py:148``) by a Jacobi-preconditioned CG on the cell graph.  The inputs are
checked as the reference checks them, with the same messages; the graph Laplacian system (cell 0
pinned, area/distance weights) is then solved matrix-free by ``csrc/pressure.hip`` with the CG
scalars on the device, and the boundary-adjacent mean is removed.  The stop test is the
reference's (SciPy's ``cg`` with ``rtol = tolerance, atol = 0``: |r| < tolerance |b|), read by the
host every ``check_every`` iterations, so the GPU may run up to ``check_every - 1`` iterations
past the point where the reference stops.

    python -m pyib_hip.postprocess RUN/final_field.npz [--out pressure.npz] [--tolerance 1e-10]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

from . import _native as N


def _finite(name, values, shape=None):
    values = np.asarray(values, dtype=np.float64)
    if shape is not None and values.shape != shape:
        raise ValueError(f"{name} must have shape {shape}, got {values.shape}")
    if not np.isfinite(values).all():
        raise ValueError(f"{name} contains non-finite values")
    return values


def check_inputs(cell_positions, face_cells, face_area, gradient):
    """The reference's input rules (ib_postprocess.py:55-78) -> (positions, pairs, areas, gradient)."""
    positions = _finite("cell_positions", cell_positions)
    count = len(positions)
    if positions.shape != (count, 3) or count < 2:
        raise ValueError("pressure recovery needs at least two 3D cells")
    gradient = _finite("gradient", gradient, positions.shape)
    pairs = np.asarray(face_cells)
    if pairs.ndim != 2 or pairs.shape[1] != 2 or pairs.dtype.kind not in "iu":
        raise ValueError("face_cells must be an integer (number_of_faces, 2) array")
    if np.any(pairs[:, 0] < 0) or np.any(pairs >= count) or np.any(pairs[:, 1] < -1):
        raise ValueError("face_cells contains invalid cell ids")
    areas = _finite("face_area", face_area, (len(pairs),))
    if np.any(areas <= 0):
        raise ValueError("face areas must be positive")
    internal = pairs[:, 1] >= 0
    d = np.linalg.norm(positions[pairs[internal, 1]] - positions[pairs[internal, 0]], axis=1)
    if not len(d) or np.any(d <= 0):
        raise ValueError("pressure graph has no valid internal faces")
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    a, b = pairs[internal, 0], pairs[internal, 1]
    adj = coo_matrix((np.ones(len(a)), (a, b)), shape=(count, count))
    if connected_components(adj, directed=False, return_labels=False) != 1:
        raise ValueError("pressure graph is disconnected; a unique gauge is unavailable")
    return positions, pairs.astype(np.int64, copy=False), areas, gradient


def recover_pressure(cell_positions, face_cells, face_area, gradient, *, tolerance=1e-10, check_every=10,
                     device="cuda"):
    """ib_postprocess.py:49-105 on the GPU -> (pressure numpy (ncell,), diagnostics dict)."""
    import torch
    from .core import dev_stream, ptr

    pos, pairs, area, grad = check_inputs(cell_positions, face_cells, face_area, gradient)
    n, m = pos.shape[0], pairs.shape[0]
    internal = pairs[:, 1] >= 0
    nnz = 2 * int(internal.sum())
    gptr, nbr, sface = np.empty(n + 1, np.int32), np.empty(nnz, np.int32), np.empty(nnz, np.int32)
    h = lambda a: a.ctypes.data_as(C.c_void_p)
    N.call("pyib_pressure_graph", C.c_int64(n), C.c_int64(m), h(np.ascontiguousarray(pairs)), h(gptr), h(nbr),
           h(sface))
    gauge_cells = np.unique(pairs[~internal, 0])
    gauge = np.zeros(n, np.uint8)
    gauge[gauge_cells if len(gauge_cells) else np.arange(n)] = 1
    dev = torch.device(device)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    keep = {"fc": t(pairs.astype(np.int32), torch.int32), "pos": t(pos, torch.float64),
            "grad": t(grad, torch.float64), "area": t(area, torch.float64), "ptr": t(gptr, torch.int32),
            "nbr": t(nbr, torch.int32), "sface": t(sface, torch.int32), "gauge": t(gauge, torch.uint8)}
    f64 = lambda k: torch.empty(k, dtype=torch.float64, device=dev)
    keep.update(w=f64(m), jmp=f64(m), dinv=f64(n), x=f64(n), r=f64(n), z=f64(n), p=f64(n), q=f64(n),
                scal=torch.zeros(8, dtype=torch.float64, device=dev),
                partials=torch.zeros(2 * N.MAX_PARTIALS, dtype=torch.float64, device=dev))
    g = N.PGraph(n, m, *(ptr(keep[k]) for k in ("fc", "pos", "grad", "area", "ptr", "nbr", "sface", "gauge", "w",
                                               "jmp", "dinv", "x", "r", "z", "p", "q", "scal", "partials")))
    st = dev_stream()
    N.call("pyib_pressure_begin", C.byref(g), st)
    maxiter = max(2000, 10 * n)
    sc = keep["scal"]
    stop = float(sc[5].item()) * tolerance * tolerance
    it = 0
    converged = stop == 0.0 or float(sc[4].item()) < stop      # b = 0: the zero pressure
    while not converged and it < maxiter:
        k = min(int(check_every), maxiter - it)
        N.call("pyib_pressure_iterate", C.byref(g), C.c_int32(k), st)
        it += k
        rr = float(sc[4].item())
        if not np.isfinite(rr):
            break
        converged = rr < stop
    if not converged:
        raise RuntimeError(f"pressure graph solve failed to converge (iterations {it})")
    out4 = torch.zeros(4, dtype=torch.float64, device=dev)
    N.call("pyib_pressure_finish", C.byref(g), ptr(out4), st)
    o = out4.cpu().numpy()
    pressure = keep["x"].cpu().numpy()
    if not np.isfinite(pressure).all():
        raise RuntimeError("pressure graph solve failed to converge (non-finite pressure)")
    return pressure, {
        "gauge": "mean boundary-adjacent cells = 0" if np.any(~internal) else "mean cells = 0",
        "gradient_projection_weighted_rms": float(np.sqrt(o[2] / o[3])),
        "iterations": it, "device": "hip",
    }


def recover_snapshot(path, *, tolerance=1e-10):
    """Whole-field pressure of a saved field (final_field.npz or a field snapshot), from the
    final inner iteration's momentum residual (ib_postprocess.py:137-149)."""
    with np.load(path, allow_pickle=False) as snap:
        for key in ("cpos", "momentum_residual", "dt", "face_cells", "face_area"):
            if key not in snap.files:
                raise ValueError(f"{Path(path).name} lacks {key}; it cannot provide pressure recovery")
        dt = float(snap["dt"])
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError("dt must be positive")
        positions = _finite("cpos", snap["cpos"])
        gradient = _finite("momentum_residual", snap["momentum_residual"], positions.shape) / dt
        return recover_pressure(positions, snap["face_cells"], snap["face_area"], gradient, tolerance=tolerance)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("field", type=Path, help="final_field.npz or a field snapshot of a run")
    ap.add_argument("--out", type=Path, default=None, help="output npz (default: <field>_pressure.npz)")
    ap.add_argument("--tolerance", type=float, default=1e-10)
    a = ap.parse_args(argv)
    try:
        pressure, diag = recover_snapshot(a.field, tolerance=a.tolerance)
    except (ValueError, RuntimeError, OSError) as exc:
        print(f"pressure recovery error: {exc}", file=sys.stderr)
        return 2
    out = a.out or a.field.with_name(a.field.stem + "_pressure.npz")
    np.savez(out, pressure=pressure, diagnostics=np.str_(json.dumps(diag, sort_keys=True)))
    print(json.dumps(dict(diag, output=str(out)), sort_keys=True))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
