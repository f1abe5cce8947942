"""Oracle restatement of the closed-sphere interior-momentum load correction.

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_interior.py``
(``_sphere_weighted_volumes`` 57-62, ``_integrate_weighted_velocity`` 65-70,
``ClosedSphereLoadCorrection`` 113-308) in float64:
F_body = -F_fluid + d/dt (interior momentum), backward difference.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np


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
    """ib_interior.py:57-62: one-cell linearly smoothed sphere indicator times volume."""
    d = np.linalg.norm(pos - center, axis=1)
    return np.clip((radius + 0.5 * width - d) / width, 0.0, 1.0) * vol


def integrate(vel, pos, wv, rho):
    """ib_interior.py:65-70."""
    lin = rho * (vel * wv[:, None]).sum(axis=0)
    ang = rho * (np.cross(pos, vel) * wv[:, None]).sum(axis=0)
    return InteriorMomenta(lin, ang, wv.sum())


class ClosedSphereLoadCorrection:
    """ib_interior.py:113-308."""

    def __init__(self, cell_positions, cell_volumes, *, reference_center, maximum_translation,
                 radius, smoothing_width, density, force_support_width=None):
        pos = np.asarray(cell_positions, float)
        vol = np.asarray(cell_volumes, float)
        c0 = np.asarray(reference_center, float)
        mt = np.asarray(maximum_translation, float)
        half = float(radius) + 0.5 * float(smoothing_width) + np.abs(mt)
        cand = np.all((pos >= c0 - half) & (pos <= c0 + half), axis=1)
        self._cand = np.nonzero(cand)[0]
        if self._cand.size == 0:
            raise ValueError("the closed-sphere indicator contains no Eulerian cells")
        self._all_pos = pos
        self._pos = pos[self._cand]
        self._vol = vol[self._cand]
        self._fidx = None
        self._fpos, self._fvol = pos, vol
        if force_support_width is not None:
            fh = float(radius) + float(force_support_width) + np.abs(mt)
            fh = fh + 8.0 * np.finfo(np.float64).eps * np.maximum(1.0, np.abs(c0) + fh)
            fm = np.all((pos >= c0 - fh) & (pos <= c0 + fh), axis=1)
            self._fidx = np.nonzero(fm)[0]
            self._fpos, self._fvol = pos[self._fidx], vol[self._fidx]
        self.radius, self.smoothing_width, self.density = float(radius), float(smoothing_width), float(density)
        self._fixed_wv = None
        if np.all(mt == 0.0):
            self._fixed_wv = weighted_volumes(self._pos, self._vol, c0, self.radius, self.smoothing_width)
        self.current = None
        self.latest_load = None

    def evaluate(self, velocity, center):
        wv = self._fixed_wv
        if wv is None:
            wv = weighted_volumes(self._pos, self._vol, np.asarray(center, float), self.radius,
                                  self.smoothing_width)
        return integrate(np.asarray(velocity, float)[self._cand], self._pos, wv, self.density)

    def initialize(self, velocity, center):
        self.current = self.evaluate(velocity, center)
        self.latest_load = None
        return self.current

    def advance(self, velocity, center, dt, force_density):
        prev = self.current
        cur = self.evaluate(velocity, center)
        lin_rate = (cur.linear_momentum - prev.linear_momentum) / float(dt)
        ang_rate = (cur.angular_momentum - prev.angular_momentum) / float(dt)
        fd = np.asarray(force_density, float)
        if self._fidx is not None:
            fd = fd[self._fidx]
        ff = (fd * self._fvol[:, None]).sum(axis=0)
        mf = (np.cross(self._fpos, fd) * self._fvol[:, None]).sum(axis=0)
        load = CorrectedBodyLoad(ff, mf, -ff + lin_rate, -mf + ang_rate, cur, lin_rate, ang_rate)
        self.current = cur
        self.latest_load = load
        return load
