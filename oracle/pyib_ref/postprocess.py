"""Offline pressure recovery of a saved field (test infrastructure: the checker of the HIP path).

Restates ``recover_pressure`` of the reference (``src/pyib/solver/ib_postprocess.py:49-105``):
the cell pressure p whose face jumps best match the saved pressure gradient g in the
area/distance-weighted least-squares sense on the cell graph,

    minimise  sum_faces  w_f (p_c2 - p_c1 - j_f)^2,   w_f = A_f / |x_c2 - x_c1|,
                                                     j_f = (g_c1 + g_c2)/2 . (x_c2 - x_c1),

solved as the graph Laplacian system with cell 0 pinned (Jacobi-preconditioned SciPy CG to a
relative residual ``tolerance``), then shifted so that the boundary-adjacent cells have zero
mean.  Input checks and messages follow the reference.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix, diags
from scipy.sparse.csgraph import connected_components
from scipy.sparse.linalg import cg


def _finite(name, values, shape=None):
    values = np.asarray(values, dtype=np.float64)
    if shape is not None and values.shape != shape:
        raise ValueError(f"{name} must have shape {shape}, got {values.shape}")
    if not np.isfinite(values).all():
        raise ValueError(f"{name} contains non-finite values")
    return values


def graph(cell_positions, face_cells, face_area, gradient):
    """(incidence, weights, jumps, internal mask) of the cell graph (ib_postprocess.py:55-92)."""
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
    first, second = pairs[internal].T
    displacement = positions[second] - positions[first]
    distance = np.linalg.norm(displacement, axis=1)
    if not len(distance) or np.any(distance <= 0):
        raise ValueError("pressure graph has no valid internal faces")
    row = np.arange(len(first))
    incidence = coo_matrix((np.r_[-np.ones(len(row)), np.ones(len(row))],
                            (np.r_[row, row], np.r_[first, second])), shape=(len(row), count)).tocsr()
    weights = areas[internal] / distance
    jumps = 0.5 * np.sum((gradient[first] + gradient[second]) * displacement, axis=1)
    return incidence, weights, jumps, internal


def recover_pressure(cell_positions, face_cells, face_area, gradient, *, tolerance=1e-10):
    """ib_postprocess.py:49-105 -> (pressure, diagnostics)."""
    incidence, weights, jumps, internal = graph(cell_positions, face_cells, face_area, gradient)
    count = incidence.shape[1]
    matrix = (incidence.T @ diags(weights) @ incidence).tocsr()
    if connected_components(matrix, directed=False, return_labels=False) != 1:
        raise ValueError("pressure graph is disconnected; a unique gauge is unavailable")
    rhs = np.asarray(incidence.T @ (weights * jumps))
    reduced = matrix[1:, 1:]
    pressure = np.zeros(count)
    pressure[1:], status = cg(reduced, rhs[1:], M=diags(1.0 / reduced.diagonal()), rtol=tolerance, atol=0.0,
                              maxiter=max(2000, 10 * count))
    if status != 0 or not np.isfinite(pressure).all():
        raise RuntimeError(f"pressure graph solve failed to converge (status {status})")
    pairs = np.asarray(face_cells)
    gauge_cells = np.unique(pairs[~internal, 0])
    if not len(gauge_cells):
        gauge_cells = np.arange(count)
    pressure -= np.mean(pressure[gauge_cells])
    mismatch = incidence @ pressure - jumps
    return pressure, {
        "gauge": "mean boundary-adjacent cells = 0" if np.any(~internal) else "mean cells = 0",
        "gradient_projection_weighted_rms": float(np.sqrt(np.sum(weights * mismatch ** 2) / weights.sum())),
    }


def recover_pressure_dense(cell_positions, face_cells, face_area, gradient):
    """The same pinned system solved densely (numpy.linalg.solve): the exact discrete solution
    for small graphs, used to pin the iterative restatement."""
    incidence, weights, jumps, internal = graph(cell_positions, face_cells, face_area, gradient)
    L = (incidence.T @ diags(weights) @ incidence).toarray()
    rhs = np.asarray(incidence.T @ (weights * jumps))
    p = np.zeros(L.shape[0])
    p[1:] = np.linalg.solve(L[1:, 1:], rhs[1:])
    pairs = np.asarray(face_cells)
    gauge = np.unique(pairs[~internal, 0])
    return p - np.mean(p[gauge if len(gauge) else np.arange(L.shape[0])])
