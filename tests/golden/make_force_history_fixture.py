"""Write the reference's D32/D40 body-force histories and harmonic fits as a fixture (build
container only: reads the reference's evidence archive, which is data).

Source: checks/pressure-validation-20260925.zip of the reference --
runs/D32|D40/history.csv (the CUDA float64 runs; columns step, time, body_force_x),
cases/D32|D40.json (motion frequency and phase) and expected.json (the published fit
coefficients of body_force_x over (4T, 6T] and the period-to-period coefficient change).

Writes tests/golden/reference_force_histories.npz (allow_pickle=False):
  <run>_step, <run>_time, <run>_body_force_x      full sampled history of the run
  <run>_omega, <run>_phase, <run>_period          from the case file
  <run>_expected_fit                              (cosine, sine, mean) from expected.json
  <run>_expected_cycle_change                     cycle_coefficient_change_over_theory_amplitude
  <run>_theory_amplitude                          |f_hat| of the unsteady Stokes drag the archive uses

Usage:  python tests/golden/make_force_history_fixture.py   (needs /root/reference)
"""
from __future__ import annotations

import csv
import io
import json
import os
import zipfile
from pathlib import Path

import numpy as np

REF = Path(os.environ.get("PYIB_REFERENCE", "/root/reference"))
OUT = Path(__file__).resolve().parent / "reference_force_histories.npz"
PREFIX = "pressure-validation-20260925/"


def main():
    z = zipfile.ZipFile(REF / "checks" / "pressure-validation-20260925.zip")
    expected = json.loads(z.read(PREFIX + "expected.json"))
    arrays = {}
    for run in ("D32", "D40"):
        case = json.loads(z.read(PREFIX + f"cases/{run}.json"))
        rows = list(csv.DictReader(io.StringIO(z.read(PREFIX + f"runs/{run}/history.csv").decode())))
        mo = case["ib"]["motion"]
        omega = 2.0 * np.pi * float(mo["frequency"])
        radius, rho = float(case["ib"]["radius"]), float(case["physics"]["density"])
        speed = float(mo["amplitude"]) * omega
        nu = speed * 2.0 * radius / float(case["physics"]["reynolds"])
        lam = (1 + 1j) * radius / np.sqrt(2.0 * nu / omega)
        f_hat = -6.0 * np.pi * rho * nu * radius * speed * (1 + lam + lam ** 2 / 9)
        bf = expected[run]["body_force"]
        arrays.update({
            f"{run}_step": np.array([int(r["step"]) for r in rows], np.int64),
            f"{run}_time": np.array([float(r["time"]) for r in rows]),
            f"{run}_body_force_x": np.array([float(r["body_force_x"]) for r in rows]),
            f"{run}_omega": np.float64(omega), f"{run}_phase": np.float64(mo["phase"]),
            f"{run}_period": np.float64(2.0 * np.pi / omega),
            f"{run}_expected_fit": np.array([bf["cosine"], bf["sine"], bf["mean"]]),
            f"{run}_expected_cycle_change": np.float64(expected[run]["cycle_coefficient_change_over_theory_amplitude"]),
            f"{run}_theory_amplitude": np.float64(abs(f_hat)),
        })
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
