"""Host-side guards (no GPU): the set-up cache key and the checkpoint source identity follow every
source that shapes the cached objects or the solution, the HIP stencil refuses a delta width it
cannot cover, and the structured layout refuses a level too large for 32-bit block offsets."""
import shutil

import numpy as np
import pytest

import pyib_hip  # noqa: F401
import pyib_hip.case as HC
import pyib_hip.mesh as PM
import pyib_hip.setup_cache as SC
from pyib_hip.immersed import HipImmersedBoundary


def test_setup_cache_key_follows_every_builder_source(tmp_path):
    here = SC.Path(SC.__file__).resolve().parent
    for name in SC._SOURCES:
        dst = tmp_path / name
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copy(here / name, dst)
    params = {"kind": "nested_box", "cells_per_diameter": 8}
    base = SC.key_for(params, tmp_path)
    assert base == SC.key_for(params, here)
    assert {"csrc/geometry.cpp", "csrc/mesh.cpp", "layout.py"} <= set(SC._SOURCES)
    for name in SC._SOURCES:
        p = tmp_path / name
        keep = p.read_bytes()
        p.write_bytes(keep + b"\n")
        assert SC.key_for(params, tmp_path) != base, name
        p.write_bytes(keep)
    assert SC.key_for(params, tmp_path) == base


def test_checkpoint_source_identity_covers_the_structured_operator():
    hashes = HC.source_hashes()
    for name in ("csrc/grid.hip", "csrc/geometry.cpp", "csrc/mesh.cpp", "layout.py", "grid.py"):
        assert name in hashes and len(hashes[name]) == 64


def test_stencil_refuses_a_kernel_wider_than_its_taps():
    mesh = PM.make_nested_mesh(0.25, PM.uniform_boxes([[-1, 1], [-1, 1], [-1, 1]]))
    HipImmersedBoundary(mesh, device="cpu")
    HipImmersedBoundary(mesh, device="cpu", hh=0.25)
    with pytest.raises(ValueError, match="hh <= dx"):
        HipImmersedBoundary(mesh, device="cpu", hh=0.5)
    with pytest.raises(ValueError):
        HipImmersedBoundary(mesh, device="cpu", hh=0.0)
    assert np.isclose(HipImmersedBoundary(mesh, device="cpu").hh, 0.25)


def test_setup_cache_write_failure_keeps_the_run(tmp_path, monkeypatch):
    """A cache that cannot be written (full disk) leaves no partial entry and returns the
    object built, with a warning."""
    from pyib_hip import setup_cache as SC
    monkeypatch.setenv("PYIB_HIP_SETUP_CACHE", str(tmp_path))

    def full(obj, path):
        path.mkdir(parents=True, exist_ok=True)
        (path / "x.npy").write_bytes(b"partial")
        raise OSError("Not enough free space")
    monkeypatch.setattr(SC, "save", full)
    with pytest.warns(RuntimeWarning, match="not stored"):
        out = SC.cached("thing", {"a": 1}, lambda: {"built": True}, dict)
    assert out == {"built": True}
    assert list(tmp_path.iterdir()) == []


def test_level_blocks_under_4gib():
    """The tiled kernels address a level's blocks with 32-bit offsets; the host refuses a level
    whose float64 cell or edge block reaches 4 GiB, and accepts the largest configured one (D40)."""
    from types import SimpleNamespace as NS
    from pyib_hip.grid import check_level_blocks
    check_level_blocks([NS(nx=240, ny=200, nz=200)])
    with pytest.raises(ValueError, match="4 GiB"):
        check_level_blocks([NS(nx=240, ny=200, nz=200), NS(nx=600, ny=600, nz=600)])
