"""Roma-Peskin 3-point immersed-boundary transfer on the GPU.

Counterpart of ``src/pyib/solver/ib_immersed.py`` (``ImmersedBoundary``):
the per-marker stencil on the finest uniform level (``index_cell``/``pos_ref``),
interpolation (Eulerian -> Lagrangian) and spreading (Lagrangian -> Eulerian).
Spreading is deterministic: the (cell, tap) pairs are radix-sorted by cell
and each cell sums its contributions in marker order, so repeated runs are
bitwise identical (the reference CuPy path scatters with atomics).

The stencil is rebuilt whenever the marker values change (the reference keys
its cache on the marker values, ``ib_immersed.py:46-61``); here the caller
bumps a version number instead of comparing arrays on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .core import dev_stream, ptr

TAPS = 27


class HipImmersedBoundary:
    def __init__(self, mesh, device="cuda", hh=None):
        idx = np.asarray(mesh.index_cell)
        self.n = np.asarray(idx.shape)
        nx, ny, nz = (int(v) for v in self.n)
        if not np.array_equal(idx.ravel(), np.arange(nx * ny * nz)):
            raise ValueError("index_cell must number the finest box in C order")
        self.pos_ref = np.asarray(mesh.pos_ref, dtype=np.float64)
        self.dx = float(np.asarray(mesh.dx_uniform)[0])
        self.hh = float(self.dx if hh is None else hh)
        # the device stencil has 27 taps (+-1 cell); the delta's support is 1.5 hh, so a wider
        # kernel (hh > dx) would need the reference's full 5^3 window (ib_immersed.py:35-41)
        if not 0.0 < self.hh <= self.dx * (1.0 + 1e-12):
            raise ValueError(f"the HIP stencil needs 0 < hh <= dx (hh={self.hh!r}, dx={self.dx!r})")
        self.ncell = int(mesh.cells.shape[0])
        self.index_cell = idx
        self.device = torch.device(device)
        self._nmark = -1
        self._version = None
        self._stencil_serial = 0        # counts stencil rebuilds (a version may be None)
        self._imp_key = None

    def _alloc(self, nmark):
        if nmark == self._nmark:
            return
        dev = self.device
        n = nmark * TAPS
        self.cell = torch.empty(n, dtype=torch.int32, device=dev)
        self.w = torch.empty(n, dtype=torch.float64, device=dev)
        self._kt = torch.empty(n, dtype=torch.int32, device=dev)
        self._vt = torch.empty(n, dtype=torch.int32, device=dev)
        self.keys = torch.empty(n, dtype=torch.int32, device=dev)
        self.vals = torch.empty(n, dtype=torch.int32, device=dev)
        nb = C.c_int64(0)
        N.call("pyib_ib_sort_temp_bytes", C.c_int64(n), C.byref(nb))
        self._temp = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=dev)
        self.vel_ib = torch.empty((nmark, 3), dtype=torch.float64, device=dev)
        self.f_lag = torch.empty((nmark, 3), dtype=torch.float64, device=dev)
        self._nmark = nmark
        self._version = None

    def stencil(self, markers, version=None):
        """Build (or reuse, when ``version`` is unchanged) the 27-tap stencil of ``markers`` (M,3) float64."""
        nmark = int(markers.shape[0])
        self._alloc(nmark)
        if version is not None and version == self._version:
            return
        nx, ny, nz = (int(v) for v in self.n)
        N.call("pyib_ib_stencil", C.c_int64(nmark), ptr(markers), float(self.pos_ref[0]), float(self.pos_ref[1]),
               float(self.pos_ref[2]), self.dx, self.hh, nx, ny, nz, ptr(self.cell), ptr(self.w), dev_stream())
        N.call("pyib_ib_sort", C.c_int64(nmark), ptr(self.cell), ptr(self._kt), ptr(self._vt), ptr(self.keys),
               ptr(self.vals), ptr(self._temp), C.c_int64(self._temp.numel()), dev_stream())
        self._version = version
        self._stencil_serial += 1

    def interp(self, vc, markers=None, version=None, out=None):
        """Eulerian -> Lagrangian (ib_immersed.py:83-93): vel = sum vc[cell] w hh^3."""
        if markers is not None:
            self.stencil(markers, version)
        out = self.vel_ib if out is None else out
        N.call("pyib_ib_interp", N.F64 if vc.dtype == torch.float64 else N.F32, C.c_int64(self._nmark),
               ptr(self.cell), ptr(self.w), self.hh ** 3, ptr(vc), ptr(out), dev_stream())
        return out

    def spread_into(self, out, flag, ds3, scale=1.0, accumulate=False):
        """Lagrangian -> Eulerian (ib_immersed.py:103-111) into ``out`` (ncell,3):
        out[c] = scale*sum (or out[c] += scale*sum) on the cells the markers touch."""
        N.call("pyib_ib_spread", N.F64 if out.dtype == torch.float64 else N.F32, C.c_int64(self._nmark),
               ptr(self.keys), ptr(self.vals), ptr(self.w), ptr(flag), ptr(ds3), float(scale),
               int(bool(accumulate)), ptr(out), dev_stream())
        return out

    def spread(self, flag, markers, ds3, version=None):
        """Fresh (ncell,3) array, as ``ImmersedBoundary.spread`` returns."""
        self.stencil(markers, version)
        out = torch.zeros((self.ncell, 3), dtype=torch.float64, device=self.device)
        return self.spread_into(out, flag, ds3)

    def implicit_lagrangian_force(self, vel_desired, dt, ds3, reg=1e-8):
        """Implicit direct forcing (ib_immersed.py:122-174) after ``interp``: f = M^-1 (u_d - interp)/dt,
        M = D S the marker coupling of the current stencil (pyib_ib_coupling), jittered by
        reg * trace(M)/n on the diagonal as the reference does; M and its inverse are rebuilt
        whenever the stencil or the marker widths change.  The dense inverse runs on the GPU."""
        key = (self._version, self._nmark, int(ds3.data_ptr()), self._stencil_serial)
        if self._imp_key != key or self._imp_key[0] is None:
            n = self._nmark
            M = torch.empty((n, n), dtype=torch.float64, device=self.device)
            N.call("pyib_ib_coupling", C.c_int64(n), ptr(self.cell), ptr(self.w), ptr(self.keys), ptr(self.vals),
                   ptr(ds3), self.hh ** 3, ptr(M), dev_stream())
            M.diagonal().add_(reg * M.diagonal().sum() / n)
            self._minv = torch.linalg.inv(M)
            self._imp_key = key
        rhs = self.lagrangian_force(vel_desired, dt)
        self.f_imp = self._minv @ rhs
        return self.f_imp

    def invalidate_implicit(self):
        """ib_immersed.py:147-150: the coupling depends on the marker positions."""
        self._imp_key = None

    def lagrangian_force(self, vel_desired, dt):
        """f_L = (u_desired - interp)/dt (ib_immersed.py:113-117), after ``interp``."""
        N.call("pyib_ib_force", C.c_int64(self._nmark), ptr(vel_desired), ptr(self.vel_ib), float(dt),
               ptr(self.f_lag), dev_stream())
        return self.f_lag
