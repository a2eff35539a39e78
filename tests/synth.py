"""Seeded synthetic light curves: the reference test suite's recipes.

* :func:`synthetic_eclipser` -- tests/synth.py:167-186 of cuPeriod (flat flux with
  periodic box dips, BJD-offset times);
* :func:`transit_curve` -- tests/test_bls_segments.py:23-29;
* :func:`two_runs` -- tests/test_bls.py:171-179 (pure noise, two runs a year apart);
* :func:`adversarial_cases` -- tests/test_bls.py:374-394 (the 16 screen-exactness cases);
* :func:`multiband_star` -- tests/test_multiband_bls_shared.py:32-57 (bands of one
  transiting star, chromatic depths 0.7x..1.3x);
* :func:`inject_transit`, :func:`template_matched_transit`, :func:`nightly_times` --
  tests/test_tls.py:36-60, 193-211, 221-224 (limb-darkened transits, nightly cadence).
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


def multiband_star(seed, *, n_band=45, n_bands=6, depth=0.02, noise=0.005, span=90.0,
                   chromatic=True):
    """``{band: (t, flux, err)}`` of one transiting star and its period."""
    rng = np.random.default_rng(seed)
    period = rng.uniform(1.5, 4.5)
    t0 = 2458000.0 + rng.uniform(0.0, period)
    bands = {}
    for b in range(n_bands):
        t = 2458000.0 + np.sort(rng.uniform(0.0, span, n_band))
        flux = 1.0 + rng.normal(0.0, noise, n_band)
        phase = np.mod(t - t0 + 0.02 * period, period) / period
        scale = 0.7 + 0.6 * b / max(1, n_bands - 1) if chromatic else 1.0
        flux[phase < 0.04] -= depth * scale
        bands[f"b{b}"] = (t, flux, np.full(n_band, noise))
    return bands, period


def inject_transit(*, n=1200, span=120.0, period=3.0, t0=2458001.0, depth=0.02,
                   dur_frac=0.05, u1=0.4, u2=0.3, noise=0.002, seed=7):
    """A flux light curve with a planted limb-darkened transit: ``(t, flux, err)``."""
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, span, n)) + 2458000.0
    phase = ((t - t0) / period) % 1.0
    x = np.where(phase > 0.5, phase - 1.0, phase)
    inside = np.abs(x) < dur_frac / 2.0
    z = np.clip(2.0 * np.abs(x) / dur_frac, 0.0, 1.0)
    mu = np.sqrt(1.0 - z * z)
    prof = 1.0 - u1 * (1.0 - mu) - u2 * (1.0 - mu) ** 2
    prof = prof / prof.max()
    flux = np.ones(n)
    flux[inside] -= depth * prof[inside]
    flux += rng.normal(0.0, noise, n)
    return t, flux, np.full(n, noise)


def template_matched_transit(width_bins, *, n=6000, period=3.0, depth=0.02, n_bins=256):
    """A noise-free transit as wide as one template: ``(t, flux)``."""
    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0.0, 120.0, n)) + 2458000.0
    q = width_bins / n_bins
    t0 = t.min() + period * 100.0 / 256.0
    phase = ((t - t0) / period) % 1.0
    x = np.where(phase > 0.5, phase - 1.0, phase)
    z = np.clip(2.0 * np.abs(x) / q, 0.0, 1.0)
    mu = np.sqrt(1.0 - z * z)
    prof = 1.0 - 0.4 * (1.0 - mu) - 0.3 * (1.0 - mu) ** 2
    flux = np.ones(n)
    inside = np.abs(x) < q / 2.0
    flux[inside] -= depth * prof[inside]
    return t, flux


def nightly_times(seed, n=370):
    """Nightly cadence over 120 nights from 2459000."""
    rng = np.random.default_rng(seed)
    night = rng.integers(0, 120, n).astype(float)
    return np.sort(night + rng.uniform(0.0, 0.28, n)) + 2459000.0
