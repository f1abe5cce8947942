"""The product's mesh geometry (C++ ``pyib_geometry``, host only, no GPU) is bitwise the
oracle's NumPy restatement of ``IBCore.__init__`` (ib_core_np.py:24-126) on 2:1 nested
meshes, including the interface centre corrections."""
import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.mesh as PM
from pyib_hip.geometry import Geometry
from oracle.pyib_ref import core as OCore

FIELDS = ("lf_face", "lf_edge", "epos", "te", "fpos", "cpos", "fe_sign_link", "nf", "Af", "rf1", "rf2", "rf",
          "vol_ci", "AoL")


@pytest.mark.parametrize("h,boxes", [(0.0625, "nested3"), (0.125, "wake"), (0.125, "concentric"),
                                     (0.5 / 24, "nested2")])
def test_native_geometry_is_bitwise_the_oracle(h, boxes):
    bx = {"nested3": PM.nested_box_boxes([1.5, 1.25, 1.25], 3), "wake": PM.wake_boxes(),
          "concentric": PM.concentric_boxes(), "nested2": PM.nested_box_boxes([1.0, 0.75, 0.75], 2)}[boxes]
    m = PM.make_nested_mesh(h, bx)
    g, o = Geometry(m), OCore.Core(m)
    for name in FIELDS:
        a, b = np.asarray(getattr(g, name)), np.asarray(getattr(o, name))
        assert a.shape == b.shape, name
        if a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint64), b.astype(np.float64).view(np.uint64)), name
        else:
            assert np.array_equal(a, b), name
    # the 2:1 corrections fired (the interface geometry is exercised)
    assert np.any(np.abs(g.AoL[g.m1] - np.median(g.AoL[g.m1])) > 1e-9)
