"""Host side of the GPU pressure recovery (no GPU): the inputs are refused as the reference
refuses them (same messages as the oracle's restatement of ib_postprocess.py:55-92), and the
cell-major adjacency of the native builder lists every internal face of a cell in face order
with the right neighbour and sign."""
import ctypes as C

import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.postprocess as PP
from pyib_hip import _native as N
from oracle.pyib_ref import postprocess as OP
from tests.test_oracle_postprocess import uniform_graph


@pytest.mark.parametrize("bad", [
    lambda p, f, a, g: (p[:1], f[:0], a[:0], g[:1]),
    lambda p, f, a, g: (p, f.astype(float), a, g),
    lambda p, f, a, g: (p, f, -a, g),
    lambda p, f, a, g: (p, np.where(f == 1, len(p), f), a, g),
    lambda p, f, a, g: (p, f, a, np.full_like(g, np.nan)),
    lambda p, f, a, g: (p, f, a, g[:, :2]),
    lambda p, f, a, g: (np.zeros_like(p), f, a, g),
])
def test_inputs_refused_like_the_reference(bad):
    pos, faces, area = uniform_graph(3)
    args = bad(pos, faces, area, np.zeros_like(pos))
    with pytest.raises(ValueError) as want:
        OP.recover_pressure(*args)
    with pytest.raises(ValueError) as got:
        PP.check_inputs(*args)
    assert str(got.value) == str(want.value)


def test_disconnected_graph_refused_before_any_device_work():
    with pytest.raises(ValueError, match="disconnected"):
        PP.recover_pressure(np.array([[0., 0., 0.], [1., 0., 0.], [3., 0., 0.]]), np.array([[0, 1], [2, -1]]),
                            np.ones(2), np.zeros((3, 3)), device="cpu")


def test_adjacency_of_the_native_builder():
    pos, faces, area = uniform_graph(4)
    n, m = len(pos), len(faces)
    internal = faces[:, 1] >= 0
    nnz = 2 * int(internal.sum())
    ptr, nbr, sface = np.empty(n + 1, np.int32), np.empty(nnz, np.int32), np.empty(nnz, np.int32)
    h = lambda a: a.ctypes.data_as(C.c_void_p)
    N.call("pyib_pressure_graph", C.c_int64(n), C.c_int64(m), h(np.ascontiguousarray(faces, np.int64)), h(ptr),
           h(nbr), h(sface))
    for c in range(n):
        want = [(int(f), int(faces[f, 1] if faces[f, 0] == c else faces[f, 0]), f if faces[f, 1] == c else -f - 1)
                for f in np.flatnonzero(internal & ((faces[:, 0] == c) | (faces[:, 1] == c)))]
        got = [(int(s if s >= 0 else -s - 1), int(o), int(s)) for o, s in zip(nbr[ptr[c]:ptr[c + 1]],
                                                                           sface[ptr[c]:ptr[c + 1]])]
        assert got == want, c
    bad = faces.copy()
    bad[0, 1] = n
    assert N.load().pyib_pressure_graph(C.c_int64(n), C.c_int64(m), h(np.ascontiguousarray(bad, np.int64)), h(ptr),
                                        h(nbr), h(sface)) != 0
