"""Build ``libpyib_hip.so`` in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", n)
       for n in ("pyib_hip.hip", "grid.hip", "adjoint.hip", "geometry.cpp", "mesh.cpp", "layout.cpp",
                 "pressure.hip", "implicit.hip")]
OUT = os.path.join(HERE, "libpyib_hip.so")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=True):
    tmp = OUT + ".tmp"
    cmd = ["hipcc", *FLAGS, "-o", tmp, *SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = [l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if r.returncode != 0:
        sys.stderr.write("\n".join(noise) + "\n")
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    if noise and verbose:
        print("\n".join(noise))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
