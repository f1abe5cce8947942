"""Closed-sphere load correction with float64 reductions on the GPU.

Counterpart of ``src/pyib/solver/ib_interior.py`` (``ClosedSphereLoadCorrection``
113-308): F_body = -F_fluid + d/dt (interior momentum), with the interior
momenta integrated over the one-cell linearly smoothed sphere indicator and
the fluid force/moment over the force-support box of the 3-point kernel.
The cell selections are made once on the host; every step runs two
deterministic float64 reductions on the device and copies back 13 numbers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _native as N
from .core import dev_stream, ptr


@dataclass
class InteriorMomenta:
    linear_momentum: Any
    angular_momentum: Any
    indicator_volume: Any


@dataclass
class CorrectedBodyLoad:
    force_fluid: Any
    moment_fluid: Any
    body_force: Any
    body_moment: Any
    interior_momenta: InteriorMomenta
    linear_momentum_rate: Any
    angular_momentum_rate: Any


def weighted_volumes(pos, vol, center, radius, width):
    """ib_interior.py:57-62 (host, for the fixed-body cache)."""
    d = np.linalg.norm(pos - center, axis=1)
    return np.clip((radius + 0.5 * width - d) / width, 0.0, 1.0) * vol


class HipLoadCorrection:
    def __init__(self, cell_positions, cell_volumes, *, reference_center, maximum_translation, radius,
                 smoothing_width, density, force_support_width=None, device="cuda"):
        pos = np.asarray(cell_positions, dtype=np.float64)
        vol = np.asarray(cell_volumes, dtype=np.float64)
        c0 = np.asarray(reference_center, dtype=np.float64)
        mt = np.asarray(maximum_translation, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3 or vol.shape != (pos.shape[0],):
            raise ValueError("cell positions must be (n,3) with matching volumes")
        if not all(np.isfinite(float(v)) and float(v) > 0.0 for v in (radius, smoothing_width, density)):
            raise ValueError("radius, smoothing_width and density must be finite and positive")
        self.device = torch.device(device)
        dev = self.device
        half = float(radius) + 0.5 * float(smoothing_width) + np.abs(mt)
        cand = np.nonzero(np.all((pos >= c0 - half) & (pos <= c0 + half), axis=1))[0]
        if cand.size == 0:
            raise ValueError("the closed-sphere indicator contains no Eulerian cells")
        self.ncell = pos.shape[0]
        self._cand = torch.from_numpy(cand.astype(np.int32)).to(dev)
        self._cpos = torch.from_numpy(np.ascontiguousarray(pos[cand])).to(dev)
        self._cvol = torch.from_numpy(np.ascontiguousarray(vol[cand])).to(dev)
        if force_support_width is not None:
            fh = float(radius) + float(force_support_width) + np.abs(mt)
            fh = fh + 8.0 * np.finfo(np.float64).eps * np.maximum(1.0, np.abs(c0) + fh)
            fidx = np.nonzero(np.all((pos >= c0 - fh) & (pos <= c0 + fh), axis=1))[0]
        else:
            fidx = np.arange(pos.shape[0])
        self._fidx = torch.from_numpy(fidx.astype(np.int32)).to(dev)
        self._fpos = torch.from_numpy(np.ascontiguousarray(pos[fidx])).to(dev)
        self._fvol = torch.from_numpy(np.ascontiguousarray(vol[fidx])).to(dev)
        self.radius, self.smoothing_width, self.density = float(radius), float(smoothing_width), float(density)
        self._fixed = bool(np.all(mt == 0.0))
        self._fwv = torch.from_numpy(weighted_volumes(pos[cand], vol[cand], c0, self.radius, self.smoothing_width)
                                     ).to(dev) if self._fixed else torch.zeros(1, dtype=torch.float64, device=dev)
        self._partials = torch.zeros(8 * N.MAX_PARTIALS, dtype=torch.float64, device=dev)
        self._out = torch.zeros(16, dtype=torch.float64, device=dev)
        self.current = None
        self.latest_load = None

    def _momenta_dev(self, vc, center):
        c = np.asarray(center, dtype=np.float64)
        N.call("pyib_interior_momenta", N.F64 if vc.dtype == torch.float64 else N.F32,
               C.c_int64(self._cand.numel()), ptr(self._cand), ptr(self._cpos), ptr(self._cvol), ptr(vc),
               float(c[0]), float(c[1]), float(c[2]), self.radius, self.smoothing_width, self.density,
               int(self._fixed), ptr(self._fwv), ptr(self._partials), ptr(self._out), dev_stream())

    def evaluate(self, velocity, center):
        """ib_interior.py:216-239; ``velocity`` is the device vc (rows beyond ncell are ignored)."""
        self._momenta_dev(velocity, center)
        o = self._out[:7].cpu().numpy()
        return InteriorMomenta(self.density * o[0:3], self.density * o[3:6], np.float64(o[6]))

    def initialize(self, velocity, center):
        self.current = self.evaluate(velocity, center)
        self.latest_load = None
        return self.current

    def restore(self, linear, angular, indicator_volume):
        lin = np.asarray(linear, dtype=np.float64)
        ang = np.asarray(angular, dtype=np.float64)
        iv = np.asarray(indicator_volume, dtype=np.float64)
        if lin.shape != (3,) or ang.shape != (3,) or iv.shape != ():
            raise ValueError("invalid closed-body checkpoint momentum shapes")
        if not (np.all(np.isfinite(lin)) and np.all(np.isfinite(ang)) and np.isfinite(iv)) or float(iv) <= 0.0:
            raise ValueError("closed-body checkpoint momenta must be finite with positive volume")
        self.current = InteriorMomenta(lin, ang, iv)
        self.latest_load = None

    def advance(self, velocity, center, dt, eulerian_force_density):
        """ib_interior.py:268-308."""
        if self.current is None:
            raise RuntimeError("closed-body correction must be initialized before advance")
        if not np.isfinite(float(dt)) or dt <= 0.0:
            raise ValueError("dt must be finite and positive")
        prev = self.current
        self._momenta_dev(velocity, center)
        F = eulerian_force_density
        N.call("pyib_fluid_load", N.F64 if F.dtype == torch.float64 else N.F32, C.c_int64(self._fidx.numel()),
               ptr(self._fidx), ptr(self._fpos), ptr(self._fvol), ptr(F), ptr(self._partials),
               ptr(self._out[8:]), dev_stream())
        o = self._out.cpu().numpy()
        cur = InteriorMomenta(self.density * o[0:3], self.density * o[3:6], np.float64(o[6]))
        lin_rate = (cur.linear_momentum - prev.linear_momentum) / float(dt)
        ang_rate = (cur.angular_momentum - prev.angular_momentum) / float(dt)
        ff, mf = o[8:11].copy(), o[11:14].copy()
        load = CorrectedBodyLoad(ff, mf, -ff + lin_rate, -mf + ang_rate, cur, lin_rate, ang_rate)
        self.current = cur
        self.latest_load = load
        return load
