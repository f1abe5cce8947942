"""Pins of the oracle's pressure recovery (oracle/pyib_ref/postprocess.py), the checker of the
GPU pressure recovery.

* The reference's own tests of recover_pressure (tests/test_ib_postprocess.py:76-85 and
  113-116): on its 9^3 uniform cell graph a quadratic pressure is recovered from its exact
  gradient up to the boundary-mean gauge (atol 3e-9, rtol 1e-8), with a projection residual
  below 1e-9, and a disconnected graph is refused.
* The iterative solve equals the dense solve of the same pinned system on small graphs.
* A planted mistake (the jump's sign reversed) fails the quadratic check.
"""
import numpy as np
import pytest

from oracle.pyib_ref import postprocess as OP


def uniform_graph(count=9, spacing=0.25):
    """The reference's test graph (tests/test_ib_postprocess.py:18-33): cell ids in C order,
    interior faces per axis, then each axis's two boundary layers with second cell -1."""
    ids = np.arange(count ** 3).reshape((count,) * 3)
    origin = -0.5 * (count - 1) * spacing
    positions = origin + spacing * np.indices(ids.shape).transpose(1, 2, 3, 0).reshape(-1, 3)
    faces = []
    for axis in range(3):
        lo, hi = [slice(None)] * 3, [slice(None)] * 3
        lo[axis], hi[axis] = slice(None, -1), slice(1, None)
        faces.append(np.column_stack((ids[tuple(lo)].ravel(), ids[tuple(hi)].ravel())))
        for side in (0, -1):
            b = [slice(None)] * 3
            b[axis] = side
            cells = ids[tuple(b)].ravel()
            faces.append(np.column_stack((cells, np.full(len(cells), -1))))
    faces = np.concatenate(faces)
    return positions, faces, np.full(len(faces), spacing ** 2)


def quadratic(x):
    x, y, z = np.asarray(x).T
    return 1.3 + 0.2 * x - 0.5 * y + 0.7 * z + 0.4 * x * x - 0.2 * y * y + 0.3 * z * z + 0.1 * x * y


def quadratic_gradient(x):
    x, y, z = np.asarray(x).T
    return np.column_stack((0.2 + 0.8 * x + 0.1 * y, -0.5 - 0.4 * y + 0.1 * x, 0.7 + 0.6 * z))


def _expected(positions, faces):
    e = quadratic(positions)
    return e - e[np.unique(faces[faces[:, 1] < 0, 0])].mean()


def test_quadratic_pressure_is_recovered():
    pos, faces, area = uniform_graph(9)
    p, diag = OP.recover_pressure(pos, faces, area, quadratic_gradient(pos))
    np.testing.assert_allclose(p, _expected(pos, faces), atol=3e-9, rtol=1e-8)
    assert abs(p[np.unique(faces[faces[:, 1] < 0, 0])].mean()) < 1e-14
    assert diag["gradient_projection_weighted_rms"] < 1e-9


def test_planted_jump_sign_error_is_caught():
    pos, faces, area = uniform_graph(5)
    p, _ = OP.recover_pressure(pos, faces, area, -quadratic_gradient(pos))
    assert not np.allclose(p, _expected(pos, faces), atol=3e-9, rtol=1e-8)


def test_disconnected_graph_is_refused():
    with pytest.raises(ValueError, match="disconnected"):
        OP.recover_pressure(np.array([[0., 0., 0.], [1., 0., 0.], [3., 0., 0.]]), np.array([[0, 1], [2, -1]]),
                            np.ones(2), np.zeros((3, 3)))


def test_iterative_equals_dense_on_a_random_gradient():
    pos, faces, area = uniform_graph(6, spacing=0.3)
    rng = np.random.default_rng(11)
    g = rng.standard_normal(pos.shape)
    p, _ = OP.recover_pressure(pos, faces, area, g, tolerance=1e-14)
    d = OP.recover_pressure_dense(pos, faces, area, g)
    np.testing.assert_allclose(p, d, rtol=0, atol=1e-11 * np.abs(d).max())


@pytest.mark.parametrize("bad,match", [
    (lambda p, f, a, g: (p[:1], f[:0], a[:0], g[:1]), "at least two"),
    (lambda p, f, a, g: (p, f.astype(float), a, g), "integer"),
    (lambda p, f, a, g: (p, f, -a, g), "positive"),
    (lambda p, f, a, g: (p, np.where(f == 1, len(p), f), a, g), "invalid cell ids"),
    (lambda p, f, a, g: (p, f, a, np.full_like(g, np.nan)), "non-finite"),
])
def test_input_errors(bad, match):
    pos, faces, area = uniform_graph(3)
    with pytest.raises(ValueError, match=match):
        OP.recover_pressure(*bad(pos, faces, area, np.zeros_like(pos)))
