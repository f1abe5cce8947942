"""The allocation-free marker sort (pyib_ib_sort_noalloc) through the C ABI.

Its outputs equal those of pyib_ib_sort with caller workspaces, bit for bit (integer work),
on two calls in a row, the second with fewer markers (the static scratch keeps the first
call's contents); they are the stable sort of the tap cells (NumPy's, on the host); spreading
from them equals spreading from pyib_ib_sort's; and marker counts beyond the static limit are
refused with a message, launching nothing."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pyib_hip  # noqa: E402,F401
from pyib_hip import _native as N  # noqa: E402
from pyib_hip.core import dev_stream, ptr  # noqa: E402


def _cells(nmark, ncell, seed):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, ncell, size=nmark * 27).astype(np.int32)
    c[rng.random(c.shape[0]) < 0.05] = -1            # taps outside the finest box
    return torch.from_numpy(c).cuda()


def _with_workspaces(cell, nmark):
    n = nmark * 27
    i32 = lambda: torch.empty(n, dtype=torch.int32, device="cuda")
    kt, vt, ks, vs = i32(), i32(), i32(), i32()
    nb = C.c_int64(0)
    N.call("pyib_ib_sort_temp_bytes", C.c_int64(n), C.byref(nb))
    temp = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device="cuda")
    N.call("pyib_ib_sort", C.c_int64(nmark), ptr(cell), ptr(kt), ptr(vt), ptr(ks), ptr(vs), ptr(temp),
           C.c_int64(temp.numel()), dev_stream())
    return ks, vs


def _no_alloc(cell, nmark):
    n = nmark * 27
    ks = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    vs = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    N.call("pyib_ib_sort_noalloc", C.c_int64(nmark), ptr(cell), ptr(ks), ptr(vs), dev_stream())
    return ks, vs


def test_no_alloc_sort_equals_the_workspace_sort_on_two_calls():
    ncell = 200_000
    for nmark, seed in ((3217, 1), (804, 2)):        # the second call is smaller
        print(f"[gpu] no-alloc sort: {nmark} markers, seed {seed}", flush=True)
        cell = _cells(nmark, ncell, seed)
        ks, vs = _with_workspaces(cell, nmark)
        ka, va = _no_alloc(cell, nmark)
        torch.cuda.synchronize()
        assert torch.equal(ka, ks) and torch.equal(va, vs)
        c = cell.cpu().numpy()
        keys = np.where(c < 0, np.int32(0x7FFFFFFF), c)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(ka.cpu().numpy(), keys[order])
        assert np.array_equal(va.cpu().numpy(), order.astype(np.int32))


def test_no_alloc_sort_feeds_the_same_spreading():
    nmark, ncell = 2000, 50_000
    cell = _cells(nmark, ncell, 3)
    rng = np.random.default_rng(4)
    w = torch.from_numpy(rng.random(nmark * 27)).cuda()
    flag = torch.from_numpy(rng.standard_normal((nmark, 3))).cuda()
    ds3 = torch.from_numpy(rng.uniform(1e-6, 2e-6, nmark)).cuda()
    outs = []
    for ks, vs in (_with_workspaces(cell, nmark), _no_alloc(cell, nmark)):
        out = torch.zeros((ncell, 3), dtype=torch.float64, device="cuda")
        N.call("pyib_ib_spread", N.F64, C.c_int64(nmark), ptr(ks), ptr(vs), ptr(w), ptr(flag), ptr(ds3), 1.0, 0,
               ptr(out), dev_stream())
        outs.append(out)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])


def test_no_alloc_sort_limits():
    top = N.IB_NOALLOC_MAX_MARKERS
    cell = _cells(top, 1_000_000, 5)
    ks, vs = _with_workspaces(cell, top)
    ka, va = _no_alloc(cell, top)                     # the largest accepted count
    torch.cuda.synchronize()
    assert torch.equal(ka, ks) and torch.equal(va, vs)
    lib = N.load()
    out = torch.empty(27, dtype=torch.int32, device="cuda")
    for bad in (top + 1, -1):
        rc = lib.pyib_ib_sort_noalloc(C.c_int64(bad), ptr(out), ptr(out), ptr(out), dev_stream())
        assert rc != 0
        assert "static scratch" in lib.pyib_last_error().decode()
    assert lib.pyib_ib_sort_noalloc(C.c_int64(0), ptr(None), ptr(None), ptr(None), dev_stream()) == 0
