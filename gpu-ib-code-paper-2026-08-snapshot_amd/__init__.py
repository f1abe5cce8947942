"""pyib_hip -- MI355X-native time-step hot path for PyIB (``--device hip``).

Python host code mirrors the reference solver's operator objects
(IBCore / Driver / ImmersedBoundary / ClosedSphereLoadCorrection) and calls
hand-written gfx950 HIP kernels through the C ABI in ``libpyib_hip.so``.
"""
__version__ = "0.1.0"
