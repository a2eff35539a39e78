"""FP64 CPU oracle of the TLS matched filter -- TEST INFRASTRUCTURE, never the product.

* :func:`tls_power` -- the numba branch of ``tls_power`` (src/cuperiod/methods/tls.py:
  598-727): ``_is_constant`` (``:185-197``), ``_flux_scale`` (``:200-213``), the weighted
  mean removal, ``_duration_bins`` (``:112-126``), ``limb_darkened_template``
  (``:81-109``), the C++ restatement of ``_numba_tls_kernel`` (``:491-595``) in
  ``tls_oracle.cpp``, the SDE and the unit and epoch corrections (``:709-727``).
* :func:`tls_bruteforce` -- the same selection from exactly summed (``math.fsum``)
  per-window correlations over the points, with no phase bins kept between windows.

Pinning: the brute force on small inputs; the template's closed form (1 at the
centre, symmetric, Claret quadratic law at the bin centres); recovery of a planted
limb-darkened transit and its unbiased depth (tests/test_tls.py:71-236); invariance
to the flux unit (``:274-288``); constant flux gives an all-zero spectrum
(``:238-258``).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from oracle import _dptr, _lib

W_EPS = 1e-300
DEN_REL_EPS = 1e-9
CONSTANT_ULPS = 64.0
OUTPUTS = ("sde", "sr", "depth", "duration", "t0")


def limb_darkened_template(n: int, u1: float, u2: float) -> np.ndarray:
    """``limb_darkened_template`` (src/cuperiod/methods/tls.py:81-109)."""
    if n <= 1:
        return np.ones(max(n, 1), dtype=np.float64)
    x = (np.arange(n, dtype=np.float64) + 0.5) / n - 0.5
    z = np.clip(2.0 * np.abs(x), 0.0, 1.0)
    mu = np.sqrt(1.0 - z * z)
    g = 1.0 - u1 * (1.0 - mu) - u2 * (1.0 - mu) ** 2
    g = np.where(z < 1.0, g, 0.0)
    peak = float(g.max())
    return g / peak if peak > 0.0 else g


def duration_bins(n_bins: int, d_min: float, d_max: float, n_dur: int) -> list[int]:
    """``_duration_bins`` (src/cuperiod/methods/tls.py:112-126)."""
    fracs = np.geomspace(d_min, d_max, n_dur)
    return sorted({min(max(int(round(f * n_bins)), 1), n_bins - 1) for f in fracs})


def is_constant(y: np.ndarray) -> bool:
    """``_is_constant`` (src/cuperiod/methods/tls.py:185-197)."""
    if y.size == 0:
        return True
    scale = float(np.max(np.abs(y)))
    return float(np.ptp(y)) <= CONSTANT_ULPS * float(np.finfo(np.float64).eps) * scale


def flux_scale(y: np.ndarray, dy: np.ndarray | None) -> float:
    """``_flux_scale`` (src/cuperiod/methods/tls.py:200-213)."""
    s = float(np.std(y)) if dy is None else float(np.median(np.abs(dy)))
    if not (math.isfinite(s) and s > 0.0):
        return 1.0
    return 2.0 ** min(max(round(math.log2(s)), -1000), 1000)


def _prepare(t, y, dy):
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    scale = flux_scale(y, dy)
    y = y / scale
    w = np.ones_like(y) if dy is None else 1.0 / (np.ascontiguousarray(dy, dtype=np.float64)
                                                   / scale) ** 2
    w_total = float(w.sum())
    y_mean = float(np.dot(w, y) / w_total)
    yw = w * (y - y_mean)
    return t, scale, w, w_total, yw


def kernel(tau, yw, w, periods, n_bins, dur_bins, templates, w_total, den_rel=DEN_REL_EPS,
           threads: int = 0):
    """The matched-filter sweep: ``{sr, depth, duration, t0}`` relative to ``tau = 0``."""
    lib = _lib()
    fn = lib.oracle_tls
    if not getattr(fn, "_typed", False):
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_int64)
        fn.restype = ctypes.c_int
        fn.argtypes = [dp, dp, dp, ctypes.c_int64, dp, ctypes.c_int64, lp, lp, dp,
                       ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                       ctypes.c_double, ctypes.c_int32, dp, dp, dp, dp]
        fn._typed = True
    tmpl = np.ascontiguousarray(np.concatenate([templates[wd] for wd in dur_bins]),
                                dtype=np.float64)
    widths = np.asarray([len(templates[wd]) for wd in dur_bins], dtype=np.int64)
    offs = np.zeros(len(dur_bins), dtype=np.int64)
    if len(dur_bins) > 1:
        offs[1:] = np.cumsum(widths[:-1])
    db = np.ascontiguousarray(dur_bins, dtype=np.int64)
    per = np.ascontiguousarray(periods, dtype=np.float64)
    n = int(per.size)
    out = {k: np.zeros(max(n, 1)) for k in ("sr", "depth", "duration", "t0")}
    lp = ctypes.POINTER(ctypes.c_int64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    yw = np.ascontiguousarray(yw, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    rc = fn(_dptr(tau), _dptr(yw), _dptr(w), int(tau.size), _dptr(per), n,
            db.ctypes.data_as(lp), offs.ctypes.data_as(lp), _dptr(tmpl), len(dur_bins),
            int(n_bins), W_EPS, float(w_total), float(den_rel), int(threads),
            _dptr(out["sr"]), _dptr(out["depth"]), _dptr(out["duration"]), _dptr(out["t0"]))
    if rc != 0:
        raise RuntimeError(f"oracle_tls failed with status {rc}")
    return {k: v[:n].copy() for k, v in out.items()}


def tls_power(t, y, dy, periods, *, n_bins=256, d_min=0.01, d_max=0.10, n_dur=5, u1=0.4,
              u2=0.3, threads: int = 0) -> dict[str, np.ndarray]:
    """The numba branch of ``tls_power`` (src/cuperiod/methods/tls.py:598-727)."""
    periods = np.ascontiguousarray(periods, dtype=np.float64)
    n = int(periods.size)
    if n == 0:
        return {k: np.zeros(0) for k in OUTPUTS}
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    t_ref = float(t.min())
    if is_constant(y):
        out = {k: np.zeros(n) for k in OUTPUTS}
        out["t0"] = np.full(n, t_ref)
        return out
    t, scale, w, w_total, yw = _prepare(t, y, dy)
    tau = t - t_ref
    db = duration_bins(n_bins, d_min, d_max, n_dur)
    templates = {wd: limb_darkened_template(wd, u1, u2) for wd in db}
    res = kernel(tau, yw, w, periods, n_bins, db, templates, w_total, threads=threads)
    sr = res["sr"]
    mean, std = float(sr.mean()), float(sr.std())
    res["sde"] = (sr - mean) / std if std > 0.0 else np.zeros_like(sr)
    res["depth"] = res["depth"] * scale
    if dy is None:
        res["sr"] = sr * (scale * scale)
    res["t0"] = res["t0"] + t_ref
    return res


def tls_bruteforce(tau, yw, w, periods, n_bins, dur_bins, templates, w_total,
                   den_rel=DEN_REL_EPS):
    """The sweep with exactly summed window correlations: for each window the points
    are gathered by phase bin (wrapping circularly) and ``num``, ``cwg``, ``den`` are
    ``math.fsum`` of their per-point terms. Same outputs as :func:`kernel`."""
    tau = np.asarray(tau, dtype=np.float64)
    ywl = np.asarray(yw, dtype=np.float64).tolist()
    wl = np.asarray(w, dtype=np.float64).tolist()
    out = {k: np.zeros(len(periods)) for k in ("sr", "depth", "duration", "t0")}
    for p, period in enumerate(np.asarray(periods, dtype=np.float64).tolist()):
        members: list[list[int]] = [[] for _ in range(n_bins)]
        for j, x in enumerate(tau.tolist()):
            q = x / period
            bb = int((q - math.floor(q)) * n_bins)
            members[min(max(bb, 0), n_bins - 1)].append(j)
        best = (0.0, 0.0, 0, 0)
        for width in dur_bins:
            g = templates[width]
            for s in range(n_bins):
                tn, tc, td = [], [], []
                for k in range(width):
                    for j in members[(s + k) % n_bins]:
                        tn.append(g[k] * ywl[j])
                        tc.append(g[k] * wl[j])
                        td.append(g[k] * g[k] * wl[j])
                nm, cw, dn = math.fsum(tn), math.fsum(tc), math.fsum(td)
                if nm < 0.0 and dn > W_EPS:
                    de = dn - cw * cw / w_total
                    if de > den_rel * dn:
                        sr = nm * nm / de
                        if sr > best[0]:
                            best = (sr, -nm / de, s, width)
        sr, depth, s, width = best
        out["sr"][p] = sr
        out["depth"][p] = depth
        out["duration"][p] = width / n_bins * period
        centre = (s + width / 2.0) / n_bins
        out["t0"][p] = (centre - math.floor(centre)) * period
    return out


__all__ = [
    "DEN_REL_EPS",
    "OUTPUTS",
    "W_EPS",
    "duration_bins",
    "flux_scale",
    "is_constant",
    "kernel",
    "limb_darkened_template",
    "tls_bruteforce",
    "tls_power",
]
