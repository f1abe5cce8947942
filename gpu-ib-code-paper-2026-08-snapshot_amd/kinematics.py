"""Prescribed rigid oscillation of the marker set (host, float64).

Same motion law as ``src/pyib/solver/ib_kinematics.py:17-40`` (``Oscillation``):
x(t) = ref + A sin(2 pi f t + phase) axis and the analytic velocity
u(t) = A 2 pi f cos(2 pi f t + phase) axis, with the axis normalised.  It is
evaluated on the host once per step; the driver uploads the (M,3) result.
"""
from __future__ import annotations

import numpy as np


class Oscillation:
    def __init__(self, ref_markers, axis=(0.0, 1.0, 0.0), A=0.1, freq=1.0, phase=0.0):
        self.ref = np.asarray(ref_markers, dtype=float)
        ax = np.asarray(axis, dtype=float)
        ax = ax / np.linalg.norm(ax)
        self.axis, self.A, self.freq, self.phase = ax, A, freq, phase
        self._w = 2.0 * np.pi * freq

    def displacement(self, t):
        return self.A * np.sin(self._w * t + self.phase) * self.axis

    def velocity(self, t):
        return self.A * self._w * np.cos(self._w * t + self.phase) * self.axis

    def __call__(self, t):
        mk = self.ref + np.asarray(self.displacement(t), dtype=float)
        vel = np.tile(np.asarray(self.velocity(t), dtype=float), (self.ref.shape[0], 1))
        return mk, vel
