"""Pins of the oracle's mesh restatement and of the product's vectorised builder.

Pins: the sha256 layout fingerprints the reference recorded for its D32 and
D40 oscillating-sphere runs (checks/pressure-validation-20260925.zip,
runs/D32|D40/metadata.json), the closed-form cell/face/edge counts, and
array-for-array equality of the two independent builders.
"""
import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.mesh as PM
from oracle.pyib_ref import case as OC
from oracle.pyib_ref import mesh as OM

D32_LAYOUT = "637889b6c63f08557ee760841b431178f4eadaa915a47344dad7282fe2daed4f"
D40_LAYOUT = "ef2ba3ca03d14e2aabd410fb3cc614180c57520e4f162258572d1d211caed13f"


def _markers(ddx):
    r, dx = 0.25, 0.5 / ddx
    n = max(12, int(round(4.0 * np.pi * r ** 2 / dx ** 2)))
    return OC.fibonacci_sphere(n, r, [0, 0, 0]), np.full(n, (4.0 * np.pi * r ** 2 * dx / n) ** (1.0 / 3.0))


@pytest.mark.parametrize("ddx,fp,counts", [
    (32, D32_LAYOUT, (1179648, 3588096, 3637824, 1229377)),
    (40, D40_LAYOUT, (2304000, 6988800, 7066320, 2381521)),
])
def test_product_builder_matches_reference_fingerprint(ddx, fp, counts):
    m = PM.make_nested_mesh(PM.DIAMETER / ddx, PM.concentric_boxes())
    assert (m.cells.shape[0], m.faces_to_cells.shape[0], m.edges_to_nodes.shape[0], m.nodes.shape[0]) == counts
    mk, w = _markers(ddx)
    assert OC.layout_fingerprint(m, mk, w) == fp


@pytest.mark.slow
def test_oracle_builder_matches_reference_fingerprint_d32():
    m = OM.make_nested_mesh(OM.DIAMETER / 32, OM.concentric_boxes())
    mk, w = _markers(32)
    assert OC.layout_fingerprint(m, mk, w) == D32_LAYOUT


@pytest.mark.parametrize("boxes", ["wake", "concentric", "nested3", "uniform"])
def test_builders_agree(boxes):
    if boxes == "wake":
        a, b = PM.make_nested_mesh(0.0625, PM.wake_boxes()), OM.make_nested_mesh(0.0625, OM.wake_boxes())
    elif boxes == "concentric":
        a, b = PM.make_nested_mesh(0.0625, PM.concentric_boxes()), OM.make_nested_mesh(0.0625, OM.concentric_boxes())
    elif boxes == "nested3":
        a = PM.make_nested_mesh(0.0625, PM.nested_box_boxes([1.5, 1.25, 1.0], 3))
        b = OM.make_nested_mesh(0.0625, OM.nested_box_boxes([1.5, 1.25, 1.0], 3))
    else:
        bb = [[-1, 1], [-0.75, 1.25], [-1, 0.5]]
        a, b = PM.make_nested_mesh(0.25, PM.uniform_boxes(bb)), OM.make_nested_mesh(0.25, OM.uniform_box(bb))
    for name in OC.LAYOUT_NAMES:
        x, y = getattr(a, name), getattr(b, name)
        assert x.dtype == y.dtype and x.shape == y.shape and np.array_equal(x, y), name


@pytest.mark.parametrize("half,levels", [([1.5, 1.25, 1.25], 2), ([3.0, 2.5, 2.5], 3), ([1.0, 1.0, 1.0], 4)])
def test_closed_form_counts(half, levels):
    m = PM.make_nested_mesh(0.0625, PM.nested_box_boxes(half, levels))
    assert PM.closed_form_counts(m.layer_shape, levels) == (m.cells.shape[0], m.faces_to_cells.shape[0],
                                                             m.edges_to_nodes.shape[0])


def test_headline_mesh_counts_closed_form():
    """3-level D32 nested box, finest 192x160x160 (the headline workload): counts by closed form."""
    shape = [(192, 160, 160)] * 3
    assert PM.closed_form_counts(shape, 3) == (13516800, 40768000, 40986624)


def test_builder_rejects_touching_boxes():
    boxes = [PM.BoxSpec.from_bounds(-1, 1, -1, 1, -1, 1), PM.BoxSpec.from_bounds(-1, 3, -2, 2, -2, 2)]
    with pytest.raises(ValueError):
        PM.make_nested_mesh(0.25, boxes)
