"""Seeded synthetic light curves: the reference test suite's recipes.

* :func:`synthetic_eclipser` -- tests/synth.py:167-186 of cuPeriod (flat flux with
  periodic box dips, BJD-offset times);
* :func:`transit_curve` -- tests/test_bls_segments.py:23-29;
* :func:`two_runs` -- tests/test_bls.py:171-179 (pure noise, two runs a year apart);
* :func:`adversarial_cases` -- tests/test_bls.py:374-394 (the 16 screen-exactness cases).
"""

from __future__ import annotations

import numpy as np


def synthetic_eclipser(
    *, n=800, span=200.0, period=2.5, depth=0.05, duration_frac=0.04,
    epoch=2458000.3, noise=0.005, base_jd=2458000.0, seed=1,
):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, span, n)) + base_jd
    phase = ((t - epoch) % period) / period
    flux = np.ones(n)
    in_transit = (phase < duration_frac) | (phase > 1.0 - duration_frac)
    flux[in_transit] -= depth
    flux += rng.normal(0.0, noise, n)
    err = np.full(n, noise)
    return t, flux, err


def transit_curve(n=800, seed=2):
    rng = np.random.default_rng(seed)
    t = 2458000.0 + np.sort(rng.uniform(0.0, 60.0, n))
    flux = 1.0 + rng.normal(0.0, 0.002, n)
    phase = np.mod(t - 2458001.3, 3.7) / 3.7
    flux[phase < 0.03] -= 0.01
    return t, flux, np.full(n, 0.002)


def two_runs(seed):
    rng = np.random.default_rng(seed)
    t = np.concatenate(
        [np.sort(rng.uniform(0.0, 3.0, 150)), np.sort(rng.uniform(365.0, 368.0, 150))]
    ) + 2458000.0
    err = rng.uniform(0.003, 0.008, t.size)
    flux = 1.0 + err * rng.standard_normal(t.size)
    return t, flux, err


def adversarial_cases():
    rng = np.random.default_rng(11)
    t, flux, err = synthetic_eclipser(n=300, span=60.0)
    n = t.size
    cases = [
        ("eclipser", t, flux, err),
        ("offset 1e6", t, flux + 1e6, err),
        ("offset 1e9", t, flux + 1e9, err),
        ("constant", t, np.ones(n), err),
        ("quantised ties", t, rng.integers(0, 3, n) * 0.01, err),
        ("wild errors", t, flux, 10.0 ** rng.uniform(-3.0, 1.0, n)),
        ("five points", t[:: n // 5][:5], flux[:: n // 5][:5], err[:: n // 5][:5]),
        ("pure noise", t, 1.0 + err * rng.standard_normal(n), err),
        ("huge errors", t, flux, np.full(n, 1e150)),
        ("both x 1e-60", t, flux * 1e-60, err * 1e-60),
        ("both x 1e60", t, flux * 1e60, err * 1e60),
    ]
    for scale in (1e-140, 1e-158, 1e-200, 1e140, 1e160):
        cases.append((f"flux x {scale:g}", t, flux * scale, err))
    return cases
