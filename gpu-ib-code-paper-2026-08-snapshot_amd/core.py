"""Device-resident mesh operators: the ``--device hip`` counterpart of IBCore.

``HipCore`` mirrors ``src/pyib/solver/ib_core_np.py`` (class ``IBCore``): the
same attributes (ncell/nface/nedge, F2C, cpos, vol_ci, Af, nf, epos, te, ...)
and the same operators (``curl``, ``rot``, ``cdiv``, ``integr_c2f``,
``interp_f2c``, ``flow_mat``), each a launch of a hand-written gfx950 kernel
through the C ABI on torch-owned device memory.  Geometry is computed once on
the host (``geometry.Geometry``); the gather tables come from ``tables``.
Host copies of the geometry stay available as NumPy arrays (``geo``) for
setup and output writers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .geometry import Geometry, diag_preconditioner
from .tables import build_tables, ghost_kinds


def dev_stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


class HipCore:
    """IBCore on the GPU.  ``precisions`` lists the coefficient sets to build:
    ``"float64"`` always, ``"float32"`` too for the mixed-precision CG."""

    def __init__(self, mesh, device="cuda", precisions=("float64",), geometry=None, tables=None):
        N.load()
        if not torch.cuda.is_available():
            raise RuntimeError("HipCore needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device)
        self.mesh = mesh
        self.geo = geo = geometry if geometry is not None else Geometry(mesh)
        self.ncell, self.nface, self.nedge = geo.ncell, geo.nface, geo.nedge
        self.F2C, self.m1 = geo.F2C, geo.m1
        self.cpos, self.fpos, self.epos, self.te = geo.cpos, geo.fpos, geo.epos, geo.te
        self.vol_ci, self.Af, self.nf, self.AoL = geo.vol_ci, geo.Af, geo.nf, geo.AoL
        self.rf, self.rf1, self.rf2 = geo.rf, geo.rf1, geo.rf2
        t = self.tables = tables if tables is not None else build_tables(geo)
        self.nghost = t.nghost
        self.bf = t.bf
        dev = self.device
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).to(dev)
        self._idx = {name: i32(getattr(t, name)) for name in
                     ("curl_ptr", "curl_idx", "rot_ptr", "rot_idx", "cf_ptr", "cf_mid", "cf_face", "cf_other", "f2c")}
        self._idx["gh_face"] = i32(t.bf)
        self._idx["gh_cell"] = i32(geo.F2C[t.bf, 0])
        self._coef = {}
        self._mesh = {}
        self.work = {}
        for prec in precisions:
            self._build_coef(prec)
        # boundary state defaults: no inlet, no Dirichlet edges, zero-gradient ghosts
        self.set_boundary(np.ones(self.nghost, np.int32), np.zeros(self.nedge, bool), np.zeros(0, np.int64),
                          np.zeros(0))

    # ------------------------------------------------------------ setup
    def _build_coef(self, prec):
        dt = torch.float64 if prec == "float64" else torch.float32
        geo = self.geo
        f = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.device, dt)
        bf = self.tables.bf
        self._coef[prec] = {
            "rf1": f(geo.rf1), "rf2": f(geo.rf2), "aol": f(geo.AoL), "vol_ci": f(geo.vol_ci),
            "gh_nf": f(geo.nf[bf]), "gh_af2": f(geo.Af[bf] ** 2),
        }
        self.work[prec] = self._alloc_work(dt)

    def _alloc_work(self, dtype):
        dev = self.device
        w = {
            "face": torch.empty(self.nface, dtype=dtype, device=dev),
            "cell_a": torch.empty(self.ncell * 3, dtype=dtype, device=dev),
            "cell_b": torch.empty(self.ncell * 3, dtype=dtype, device=dev),
            "partials": torch.zeros(8 * N.MAX_PARTIALS, dtype=torch.float64, device=dev),
        }
        w["struct"] = N.Work(ptr(w["face"]), ptr(w["cell_a"]), ptr(w["cell_b"]), ptr(w["partials"]))
        return w

    def set_boundary(self, gh_kind, dir_mask, inlet_edge, inlet_val):
        """Install ghost rules, Dirichlet mask and inlet stream values."""
        dev = self.device
        self._idx["gh_kind"] = torch.from_numpy(np.ascontiguousarray(gh_kind, np.int32)).to(dev)
        self.dir_mask_host = np.asarray(dir_mask, bool)
        self._dir = torch.from_numpy(self.dir_mask_host.astype(np.uint8)).to(dev)
        self._inlet_edge = torch.from_numpy(np.ascontiguousarray(inlet_edge, np.int32)).to(dev)
        self._inlet_val = torch.from_numpy(np.ascontiguousarray(inlet_val, np.float64)).to(dev)
        for prec in self._coef:
            self._mesh[prec] = self._make_struct(prec)

    def _make_struct(self, prec):
        c, ix = self._coef[prec], self._idx
        m = N.Mesh()
        m.dtype = N.F64 if prec == "float64" else N.F32
        m.ncell, m.nface, m.nedge, m.nghost = self.ncell, self.nface, self.nedge, self.nghost
        for name in ("curl_ptr", "curl_idx", "rot_ptr", "rot_idx", "cf_ptr", "cf_mid", "cf_face", "cf_other", "f2c",
                     "gh_face", "gh_cell", "gh_kind"):
            setattr(m, name, ptr(ix[name]))
        for name in ("rf1", "rf2", "aol", "vol_ci", "gh_nf", "gh_af2"):
            setattr(m, name, ptr(c[name]))
        m.ninlet = int(self._inlet_edge.numel())
        m.inlet_edge = ptr(self._inlet_edge)
        m.inlet_val = ptr(self._inlet_val)
        m.dir_mask = ptr(self._dir)
        return m

    def struct(self, prec="float64"):
        return C.byref(self._mesh[prec])

    def work_struct(self, prec="float64"):
        return C.byref(self.work[prec]["struct"])

    def preconditioner(self, dir_mask):
        return diag_preconditioner(self.geo, dir_mask)

    def tensor(self, a, dtype=torch.float64):
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device, dtype)

    # ----------------------------------------------- IBCore operators
    def _check(self, t, n, dtype=torch.float64):
        if not (t.is_cuda and t.is_contiguous() and t.numel() == n and t.dtype == dtype):
            raise ValueError(f"expected a contiguous {dtype} device tensor with {n} elements")

    def curl(self, edge, out=None):
        """ib_core_np.py:133-134 (edge -> face)."""
        self._check(edge, self.nedge, edge.dtype)
        out = torch.empty(self.nface, dtype=edge.dtype, device=self.device) if out is None else out
        N.call("pyib_curl", self.struct(_prec(edge)), ptr(edge), ptr(out), dev_stream())
        return out

    def rot(self, face, out=None):
        """ib_core_np.py:129-131 (face -> edge)."""
        self._check(face, self.nface, face.dtype)
        out = torch.empty(self.nedge, dtype=face.dtype, device=self.device) if out is None else out
        N.call("pyib_rot", self.struct(_prec(face)), ptr(face), ptr(out), dev_stream())
        return out

    def cdiv(self, face, out=None):
        """ib_core_np.py:136-139 (face -> cell)."""
        self._check(face, self.nface, face.dtype)
        out = torch.empty(self.ncell, dtype=face.dtype, device=self.device) if out is None else out
        N.call("pyib_cdiv", self.struct(_prec(face)), ptr(face), ptr(out), dev_stream())
        return out

    def integr_c2f(self, cellvec, out=None):
        """ib_core_np.py:144-148 (cell vector -> face); cellvec may carry ghost rows after ncell."""
        if not (cellvec.is_cuda and cellvec.is_contiguous() and cellvec.shape[-1] == 3):
            raise ValueError("integr_c2f expects a contiguous (n,3) device tensor")
        out = torch.empty(self.nface, dtype=cellvec.dtype, device=self.device) if out is None else out
        N.call("pyib_integr_c2f", self.struct(_prec(cellvec)), ptr(cellvec), ptr(out), dev_stream())
        return out

    def interp_f2c(self, face, out=None):
        """ib_core_np.py:150-157 (face -> cell vector)."""
        self._check(face, self.nface, face.dtype)
        out = torch.empty((self.ncell, 3), dtype=face.dtype, device=self.device) if out is None else out
        N.call("pyib_interp_f2c", self.struct(_prec(face)), ptr(face), ptr(out), dev_stream())
        return out

    def flow_mat(self, pp, alpd=0.0, visc=0.0, rou=1.0, dt=0.0, out=None, mask_dir=False):
        """ib_core_np.py:168-175 (matrix-free elliptic operator).  With alpd == 0 the
        viscous block is the identity (rou=1, visc=0)."""
        self._check(pp, self.nedge, pp.dtype)
        prec = _prec(pp)
        out = torch.empty(self.nedge, dtype=pp.dtype, device=self.device) if out is None else out
        if alpd == 0.0:
            rou, visc, dt = 1.0, 0.0, 0.0
        N.call("pyib_flow_mat", self.struct(prec), self.work_struct(prec), float(rou), float(visc), 0.5 * float(dt),
               ptr(pp), ptr(out), int(bool(mask_dir)), C.c_void_p(0), dev_stream())
        return out


def _prec(t):
    if t.dtype == torch.float64:
        return "float64"
    if t.dtype == torch.float32:
        return "float32"
    raise ValueError(f"unsupported dtype {t.dtype}")
