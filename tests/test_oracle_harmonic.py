"""Pins of the oracle's harmonic fit (oracle/pyib_ref/case.py harmonic_fit), the definition of
the headline force-harmonic metric (tools/check_pressure_validation.py:32-38 of the reference).

* The reference's own evidence: refitting the CUDA float64 D32 and D40 body-force histories
  over (4T, 6T] reproduces the coefficients the reference published (expected.json), to the
  tolerance its audit uses (rtol 1e-9, atol 1e-12; check_pressure_validation.py:45), and the
  period-to-period coefficient change over (4T, 5T] and (5T, 6T] reproduces the published one.
* Closed forms: an exact a cos + b sin + c is recovered to round-off, and a fit of noisy data
  equals the solution of the 3x3 normal equations; a rank-deficient window raises.
* A planted mistake (phase sign flipped) fails the published-coefficient check.
"""
import os

import numpy as np
import pytest

from oracle.pyib_ref.case import harmonic_fit

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def hist():
    with np.load(os.path.join(GOLDEN, "reference_force_histories.npz"), allow_pickle=False) as g:
        return {k: g[k] for k in g.files}


def _window(h, run, lo, hi):
    T, t = float(h[f"{run}_period"]), h[f"{run}_time"]
    return (t > lo * T + 1e-10 * T) & (t <= hi * T + 1e-10 * T)


@pytest.mark.parametrize("run", ["D32", "D40"])
def test_fit_reproduces_published_coefficients(hist, run):
    use = _window(hist, run, 4, 6)
    assert use.sum() > 100
    c = harmonic_fit(hist[f"{run}_time"][use], hist[f"{run}_body_force_x"][use], float(hist[f"{run}_omega"]),
                     float(hist[f"{run}_phase"]))
    np.testing.assert_allclose(c, hist[f"{run}_expected_fit"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("run", ["D32", "D40"])
def test_cycle_change_reproduces_published_value(hist, run):
    t, f = hist[f"{run}_time"], hist[f"{run}_body_force_x"]
    w, ph = float(hist[f"{run}_omega"]), float(hist[f"{run}_phase"])
    prev = harmonic_fit(t[_window(hist, run, 4, 5)], f[_window(hist, run, 4, 5)], w, ph)
    last = harmonic_fit(t[_window(hist, run, 5, 6)], f[_window(hist, run, 5, 6)], w, ph)
    change = float(np.linalg.norm(last - prev) / float(hist[f"{run}_theory_amplitude"]))
    np.testing.assert_allclose(change, float(hist[f"{run}_expected_cycle_change"]), rtol=1e-9, atol=1e-12)


def test_planted_phase_sign_error_is_caught(hist):
    use = _window(hist, "D32", 4, 6)
    c = harmonic_fit(hist["D32_time"][use], hist["D32_body_force_x"][use], float(hist["D32_omega"]),
                     -float(hist["D32_phase"]))
    assert not np.allclose(c, hist["D32_expected_fit"], rtol=1e-9, atol=1e-12)


def test_exact_harmonic_is_recovered():
    w, ph = 3.6, -0.5 * np.pi
    t = np.linspace(0.0, 2.0 * np.pi / w, 97)[1:]
    a, b, c = -0.0123, 0.0101, 1.7e-5
    y = a * np.cos(w * t + ph) + b * np.sin(w * t + ph) + c
    np.testing.assert_allclose(harmonic_fit(t, y, w, ph), [a, b, c], rtol=1e-12, atol=1e-17)


def test_noisy_fit_equals_normal_equations():
    rng = np.random.default_rng(7)
    w, ph = 2.0, 0.3
    t = np.sort(rng.uniform(0.0, 7.0, 200))
    y = rng.standard_normal(200)
    A = np.column_stack([np.cos(w * t + ph), np.sin(w * t + ph), np.ones_like(t)])
    want = np.linalg.solve(A.T @ A, A.T @ y)
    np.testing.assert_allclose(harmonic_fit(t, y, w, ph), want, rtol=1e-10, atol=1e-14)


def test_rank_deficient_window_raises():
    with pytest.raises(ValueError):
        harmonic_fit(np.array([0.0, 0.0, 0.0]), np.array([1.0, 2.0, 3.0]), 1.0, 0.0)
