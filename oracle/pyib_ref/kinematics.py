"""Oracle restatement of the prescribed rigid oscillation.

TEST INFRASTRUCTURE ONLY.  Follows ``src/pyib/solver/ib_kinematics.py:17-40``
(``RigidMotion``/``Oscillation``): x(t) = ref + A sin(2 pi f t + phase) axis,
u(t) = A 2 pi f cos(2 pi f t + phase) axis, axis normalised.
"""
from __future__ import annotations

import numpy as np


class Oscillation:
    def __init__(self, ref_markers, axis=(0.0, 1.0, 0.0), A=0.1, freq=1.0, phase=0.0):
        self.ref = np.asarray(ref_markers, dtype=float)
        ax = np.asarray(axis, dtype=float)
        self.axis = ax / np.linalg.norm(ax)
        self.A, self.freq, self.phase = A, freq, phase
        self.w = 2.0 * np.pi * freq

    def __call__(self, t):
        mk = self.ref + np.asarray(self.A * np.sin(self.w * t + self.phase) * self.axis, dtype=float)
        vel = np.tile(np.asarray(self.A * self.w * np.cos(self.w * t + self.phase) * self.axis, dtype=float),
                      (self.ref.shape[0], 1))
        return mk, vel
