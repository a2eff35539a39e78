"""Benchmark: batched plain BLS on one MI355X, with the oracle and the other legs beside it.

Headline workload (the survey batch, ``cuperiod_amd.workloads.survey_batch``): 4,096
curves, N log-uniform in [100, 5,000], baselines 50-2,000 d, 200,000 linspace frequencies
per curve in [1/baseline, 2] /d, 6 period-scaled durations per log segment. A step is one
replay of the captured batched search (descriptor upload, curve preparation, the main
kernel's launch classes, the reference rounding of each segment's peak, per-curve
summaries) with the inputs resident in HBM. W untimed warm-up steps, then K timed steps
between two device synchronisations; ``value`` = curves/s. One JSON line is printed.

``roofline``: the main kernel runs on the vector ALUs (no MFMA, HBM far from busy), so the
bound is ``valu``: algorithmic FLOP per launch (4 FP64 per point-fold, 8 FP32 per box
evaluation) over its mean HIP-event duration, against the vector FMA peak (78.6 TFLOP/s:
256 CUs x 4 SIMDs x 16 lanes x 2 FLOP x 2.4 GHz, the published FP64 vector peak, at which
non-packed FP32 also issues). ``issue`` adds the measured VALU instruction rate against
the issue limit (256 x 4 x 16 lanes x 2.4 GHz), from rocprofv3 SQ_INSTS_VALU counts of the
same workload.

Legs beside the headline (each skippable):
* ``parity``: the first ``--cpu-curves`` curves (64) on the engine and the FP64 oracle at
  full grids: max relative power error, best periods identical, and every differing best
  period proven an exact equal-point tie of the oracle's own maximum;
* ``cpu_baseline``: the oracle's time on that sample (all host threads), in curves/s,
  point-folds/s and box-evaluations/s, extrapolated to the whole batch; plus a
  calibration run on the published single-curve shape against the published numba rate;
* ``end_to_end``: ``batch_periodograms`` on the whole batch (host checks, packing,
  uploads, search, peak candidates, rows) in curves/s;
* ``long_baseline``: the long-baseline batch (256 curves x 100,000 points x 4,000,000
  frequencies, 16 durations) timed on a seeded subset and extrapolated in sum(N x F), with
  the precision modes' errors against the oracle;
* ``learned``: the learned proposer + refinement on the learned batch (16,384 gapped
  curves x 200,000 frequencies) against the plain search: speed-up and agreement of the
  best period;
* ``next_rows``: joint multi-band BLS and TLS, each one batched launch beside its oracle.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("light curves/s, 200k-freq BLS, 1 GPU: learned+refine vs plain, best-period "
          "match %")
PEAK_VALU_FLOPS = 78.6e12
VALU_ISSUE = 256 * 4 * 16 * 2.4e9
PEAK_LDS = 256 * 128 * 2.4e9
COPY_BYTES = 256 << 20
#: Published single-curve numba BLS rate on 16 cores / 32 threads (891 points x 76,124
#: periods in 39.5 ms, cuPeriod benchmarks/REPORT.md:206).
PUBLISHED_NUMBA_FOLDS_PER_S = 891 * 76124 / 0.039467
PUBLISHED_NUMBA_CORES = 16
#: Per step of the survey batch (4,096 curves x 200,000 frequencies): HBM bytes of
#: bls_main_kernel (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by a plain copy's
#: reading) and its VALU wave instructions (--pmc SQ_INSTS_VALU), summed over the launch
#: classes; the commands are in profiles/bls_main_kernel_survey_batch.md.
MEASURED = {(4096, 200_000): {"traffic": 58.90e9, "valu_wave_instr": 2.524e12}}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def count_work(plans, n_points, oversample) -> tuple[float, float]:
    """Exact point-folds and box evaluations of a search over ``plans``."""
    from cuperiod_amd.methods._bls_core import _duration_bins

    folds = boxes = 0.0
    for plan, n in zip(plans, n_points):
        for sg in plan.segments:
            sub = type(plan)(segments=[sg], n_periods=sg.n_periods)
            periods = sub.periods()
            bd = float(sg.durations.min()) / oversample
            width = int(np.ceil(sg.p_max / bd)) + oversample + 1
            nb = np.minimum(np.ceil(periods / bd).astype(np.int64) + oversample, width - 1)
            folds += float(n) * sg.n_periods
            for d in _duration_bins(sg.durations, bd):
                m = nb[nb > d]
                boxes += float((m - d + 1).sum())
    return folds, boxes


def _kernel_ms(ctx) -> float:
    ms = ctypes.c_double()
    ctx.check(ctx.lib.cpb_main_kernel_ms(ctx.handle, ctypes.byref(ms)), "timing")
    return ms.value


def _oracle_curve(t, y, dy, segs, settings, threads):
    import oracle

    parts = [oracle.bls_power(t, y, dy, p, d, settings.bins_per_duration,
                              objective=settings.objective, threads=threads)
             for p, d in segs]
    return {k: np.concatenate([p[k] for p in parts]) for k in ("period", *oracle.OUTPUT_NAMES)}


def parity_and_cpu_leg(wl, settings, pb, n: int) -> tuple[dict, dict]:
    """The oracle on the first ``n`` curves: its time (the CPU baseline) and the engine's
    agreement with it on the same curves and grids."""
    import oracle
    from cuperiod_amd.batch.engine import bls_batch
    from cuperiod_amd.core.result import Periodogram
    from cuperiod_amd.methods.bls import _grid_segments
    from oracle.ties import BoxInputs, prove_cross_period_tie

    threads = oracle.default_threads()
    segs = [_grid_segments(wl.grids[c].as_gridspec(), settings) for c in range(n)]
    t0, y0, e0 = wl.curves[0]
    oracle.bls_power(t0[:50], y0[:50], e0[:50], np.array([1.0, 2.0]), np.array([0.1]), 10)
    start = time.perf_counter()
    refs = [_oracle_curve(*wl.curves[c], segs[c], settings, threads) for c in range(n)]
    elapsed = time.perf_counter() - start
    eng = bls_batch(wl.curves[:n], wl.grids[:n], settings, spectra=True)
    worst = 0.0
    identical = ties = 0
    tie_ulps = []
    for c in range(n):
        ref, got = refs[c], eng.spectra[c]
        finite = np.isfinite(ref["power"]) & np.isfinite(got["power"])
        rel = np.abs(got["power"][finite] - ref["power"][finite]) / np.maximum(
            np.abs(ref["power"][finite]), 1e-300)
        worst = max(worst, float(rel.max()) if rel.size else 0.0)
        rpow = np.nan_to_num(ref["power"], neginf=0.0)
        pg = Periodogram.from_spectrum(method="BLS", backend="oracle",
                                       frequency=1.0 / ref["period"], power=rpow,
                                       objective_sense="max", n_samples=1, baseline=1.0)
        ref_best = pg.best_period()
        if eng.best_period[c] == ref_best:
            identical += 1
        else:
            gi = int(eng.best_index[c])
            ri = int(np.flatnonzero(ref["period"] == ref_best)[0])
            inputs = BoxInputs.for_specs(wl.curves[c][0], segs[c], settings.bins_per_duration)
            tie_ulps.append(prove_cross_period_tie(ref, inputs, gi, ri, name=f"curve {c}"))
            ties += 1
    folds, boxes = count_work(pb.plans[:n], [wl.curves[c][0].size for c in range(n)],
                              settings.bins_per_duration)
    parity = {
        "sample": f"the first {n} curves of the batch, every frequency of their grids",
        "max_rel_power_error": worst,
        "power_tolerance": 1e-6,
        "best_period_identical": identical / n,
        "best_period_equal_point_ties": ties,
        "tie_oracle_power_ulps": tie_ulps,
        "best_period_match": (identical + ties) / n,
        "tie_rule": ("a differing best period counts only if both periods' boxes hold the "
                     "same multiset of points (equal exact statistics) and the oracle's "
                     "own powers there agree to FP64 summation-order rounding (1e-12 "
                     "relative); tie_oracle_power_ulps gives their distance"),
    }
    cpu = {
        "value": round(n / elapsed, 5),
        "unit": "light curves/s",
        "cores": int(threads),
        "kind": "port",
        "sample": f"the first {n} curves of the same batch, same grids and settings, one "
                  f"timed pass ({elapsed:.1f} s)",
        "point_folds_per_s": folds / elapsed,
        "box_evals_per_s": boxes / elapsed,
        "sample_point_folds": folds,
    }
    return parity, cpu


def calibration_leg() -> dict:
    """The oracle and the engine on the published single-curve shape."""
    import oracle
    from cuperiod_amd.core.lightcurve import LightCurve
    from cuperiod_amd.methods.bls import BLSMethod, _grid_segments
    from cuperiod_amd.workloads import published_curve

    t, y, dy, settings = published_curve()
    lc = LightCurve.from_arrays(t, y, dy, domain="flux")
    method = BLSMethod()
    grid = method.default_grid(lc, settings)
    segs = _grid_segments(grid, settings)
    threads = oracle.default_threads()
    runs = []
    _oracle_curve(t, y, dy, segs, settings, threads)  # warm-up
    for _ in range(3):
        s0 = time.perf_counter()
        _oracle_curve(t, y, dy, segs, settings, threads)
        runs.append(time.perf_counter() - s0)
    cpu_s = min(runs)
    method.power(grid, lc, settings, "hip")  # warm-up
    import torch

    runs = []
    for _ in range(3):
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        method.power(grid, lc, settings, "hip")
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - s0)
    gpu_s = min(runs)
    folds = float(t.size) * float(grid.values.size)
    rate = folds / cpu_s
    expected = PUBLISHED_NUMBA_FOLDS_PER_S * threads / PUBLISHED_NUMBA_CORES
    return {
        "workload": f"N={t.size} over 2,175 d, {grid.values.size} default-grid periods in "
                    "0.5-4 d (grid_oversample 1): the published single-curve shape",
        "oracle_s": round(cpu_s, 5),
        "oracle_point_folds_per_s": rate,
        "published_numba_point_folds_per_s": PUBLISHED_NUMBA_FOLDS_PER_S,
        "published_numba_cores": PUBLISHED_NUMBA_CORES,
        "published_scaled_to_these_cores": expected,
        "oracle_over_published_scaled": round(rate / expected, 3),
        "engine_single_curve_s": round(gpu_s, 5),
        "published_numba_s": 0.039467,
        "published_cupy_s": 0.0898,
        "method": "warm-up, then best of 3 (benchmarks/benchmark.py:62-64)",
        "host_cpu": _cpu_model(),
        "published_host_cpu": "AMD Ryzen 9 9950X3D, 16 cores / 32 threads",
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def fp64_error_leg(parity: dict | None, long_baseline: dict | None) -> dict:
    """Largest relative power error against the FP64 oracle on each single-GPU workload:
    the small curve and the gapped curve searched here, the survey batch's parity sample
    and the long-baseline precision study taken from their legs."""
    import oracle
    from cuperiod_amd.methods import _bls_core as core
    from cuperiod_amd.workloads import gapped_curve, small_curve

    def rel_error(got, ref) -> float:
        ok = np.isfinite(got) & np.isfinite(ref)
        if not ok.any():
            return 0.0
        return float(np.max(np.abs(got[ok] - ref[ok]) / np.maximum(np.abs(ref[ok]), 1e-300)))

    found = {}
    wl = small_curve()
    t, y, dy = wl.curves[0]
    periods = wl.grids[0].periods()
    got = core.bls_power(t, y, dy, periods, wl.durations, 10).power
    ref = oracle.bls_power(t, y, dy, periods, wl.durations, 10)["power"]
    best = lambda p: int(np.argmax(np.nan_to_num(p, nan=-np.inf, neginf=-np.inf)))  # noqa: E731
    found["small curve: N=2,000 over 100 d, 50,000 frequencies, 4 durations"] = {
        "max_rel_power_error": rel_error(got, ref), "periods_checked": int(periods.size),
        "best_period_identical": bool(periods[best(got)] == periods[best(ref)])}
    wl = gapped_curve()
    t, y, dy = wl.curves[0]
    periods = wl.grids[0].periods()
    got = core.bls_power(t, y, dy, periods, wl.durations, 10).power
    k = best(got)
    rng = np.random.default_rng(77)
    idx = np.unique(np.concatenate([rng.choice(periods.size, 3000, replace=False),
                                    np.arange(max(0, k - 200), min(periods.size, k + 200))]))
    ref = oracle.bls_power(t, y, dy, periods[idx], wl.durations, 10)["power"]
    found["gapped curve: N~14,000 over 1,000 d, 1,000,000 frequencies, 8 durations"] = {
        "max_rel_power_error": rel_error(got[idx], ref), "periods_checked": int(idx.size),
        "best_period_identical": bool(periods[idx][best(ref)] == periods[k])}
    if parity:
        found["survey batch: the parity sample"] = {
            "max_rel_power_error": parity["max_rel_power_error"],
            "best_period_match": parity["best_period_match"]}
    if long_baseline and long_baseline.get("precision"):
        fx = long_baseline["precision"]["modes"]["fixed64"]
        found["long-baseline curve: N=100,000, 8,000 frequencies of 4,000,000"] = {
            "max_rel_power_error": fx["max_rel_power_error"],
            "best_period_identical": fx["best_period_identical"]}
    return found


def end_to_end_leg(wl, settings, chunk: int) -> dict:
    """``batch_periodograms`` on the whole batch: checks, packing, uploads, search, peak
    candidates and rows, with each curve's 200,000-frequency linspace grid."""
    import torch

    from cuperiod_amd.batch.engine import LinspaceGrid
    from cuperiod_amd.batch.runner import batch_periodograms
    from cuperiod_amd.core.lightcurve import LightCurve

    items = [(f"c{c:05d}", LightCurve.from_arrays(t, y, dy, domain="flux"))
             for c, (t, y, dy) in enumerate(wl.curves)]
    n_freq = wl.grids[0].num

    def rule(lc):
        return LinspaceGrid(1.0 / lc.baseline, 2.0, n_freq)

    # untimed warm-up on one chunk of the same size (device buffers of that size exist)
    batch_periodograms(items[:chunk], "BLS", grid=rule, settings=settings, chunk_size=chunk)
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    res = batch_periodograms(items, "BLS", grid=rule, settings=settings, chunk_size=chunk)
    torch.cuda.synchronize()
    el = time.perf_counter() - s0
    return {
        "workload": f"batch_periodograms, {len(items)} survey curves, each on linspace(1/"
                    f"baseline, 2, {n_freq}) /d, {settings.n_durations} segment durations, "
                    f"chunks of {chunk}, rows with 10 peaks",
        "value": round(len(items) / el, 3),
        "unit": "light curves/s",
        "wall_s": round(el, 3),
        "rows": res.n_done,
        "failed": res.n_failed,
        "includes": "per-curve checks, host preparation and packing, H2D, the search, "
                    "peak candidates and their D2H, host peak selection into rows",
        "warmup": f"one untimed call on the first {chunk} curves",
    }


def long_baseline_leg(n_timed: int, cpu: bool) -> dict:
    """The long-baseline batch: a seeded subset of its 256 curves timed at the full
    4,000,000-frequency grid and extrapolated in sum(N x F); the precision modes against
    the oracle on windows of one curve's grid."""
    import torch

    from cuperiod_amd import _lib
    from cuperiod_amd.batch import engine
    from cuperiod_amd.core.config import BLSSettings
    from cuperiod_amd.workloads import long_baseline_batch

    settings = BLSSettings(n_durations=16)
    wl = long_baseline_batch(n_curves=n_timed)
    ctx = _lib.context(0)
    ctx.lib.cpb_set_timing(ctx.handle, 1)
    pb = engine.pack(wl.curves, wl.grids, settings)
    out, summary = engine.run_packed(pb)  # warm-up
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    out, summary = engine.run_packed(pb, out=out, summary=summary)
    torch.cuda.synchronize()
    el = time.perf_counter() - s0
    k_ms = _kernel_ms(ctx)
    res = engine.finish(pb, out, summary)
    folds = float(sum(c[0].size for c in wl.curves)) * float(wl.grids[0].num)
    full_folds = 256.0 * 100_000 * float(wl.grids[0].num)
    leg = {
        "workload": f"long-baseline batch: {n_timed} of its 256 curves timed (N=100,000 over "
                    "10,000 d offset by 2.45e6 d, 4,000,000 frequencies in [1e-4, 10] /d, "
                    "16 segment durations); extrapolated to 256 curves in sum(N x F)",
        "value": round(n_timed / el, 4),
        "unit": "light curves/s",
        "timed_curves": n_timed,
        "wall_s": round(el, 3),
        "main_kernel_ms": round(k_ms, 3),
        "point_folds_per_s": folds / el,
        "extrapolated_batch_s": round(el * full_folds / folds, 2),
        "best_periods": [float(p) for p in res.best_period],
        "injected_period": 0.7,
        "precision": None,
    }
    if cpu:
        leg["precision"] = precision_study()
    return leg


def precision_study(windows: int = 4000) -> dict:
    """The engine's precision modes against the FP64 oracle on one long-baseline curve:
    two windows of its grid (the top at 10 /d, phases ~1e5 cycles; around 1/0.7 d)."""
    import oracle
    from cuperiod_amd.batch import engine
    from cuperiod_amd.core.config import BLSSettings
    from cuperiod_amd.workloads import long_baseline_batch

    base = BLSSettings(n_durations=16)
    wl = long_baseline_batch(n_curves=1)
    grid = wl.grids[0]
    full = engine._linspace_plan(0, grid, base)
    step = (grid.stop - grid.start) / (grid.num - 1)
    k_tr = grid.num - 1 - int(round((1.0 / 0.7 - grid.start) / step))
    plan = engine.window_plan(full, [(0, windows), (k_tr - windows // 2, k_tr + windows // 2)])
    t, y, dy = wl.curves[0]
    threads = oracle.default_threads()
    parts = []
    s0 = time.perf_counter()
    for sg in plan.segments:
        sub = engine.CurvePlan(segments=[sg], n_periods=sg.n_periods)
        parts.append(oracle.bls_power(t, y, dy, sub.periods(), sg.durations,
                                      base.bins_per_duration, threads=threads))
    cpu_s = time.perf_counter() - s0
    ref = {k: np.concatenate([r[k] for r in parts]) for k in parts[0]}
    ok = np.isfinite(ref["power"]) & (np.abs(ref["power"]) > 0)
    best_ref = ref["period"][int(np.argmax(np.where(ok, ref["power"], -np.inf)))]
    cs = engine.pack_curves(wl.curves)
    modes = {}
    for acc in ("fixed64", "kahan32", "fp32"):
        settings = BLSSettings(n_durations=16, accumulation=acc)
        pb = engine.pack_plans(cs, [plan], settings)
        res = engine.finish(pb, *engine.run_packed(pb), spectra=True)
        p = res.spectra[0]["power"]
        rel = np.abs(p[ok] - ref["power"][ok]) / np.abs(ref["power"][ok])
        best = plan.periods()[int(np.argmax(np.where(ok, p, -np.inf)))]
        modes[acc] = {"max_rel_power_error": float(rel.max()),
                      "median_rel_power_error": float(np.median(rel)),
                      "best_period_identical": bool(best == best_ref)}
    folds = float(t.size) * float(plan.n_periods)
    return {"curve": "long-baseline curve 0 (N=100,000, t offset 2.45e6 d)",
            "frequencies": int(plan.n_periods), "oracle_best_period": float(best_ref),
            "modes": modes,
            "cpu_oracle": {"seconds": round(cpu_s, 3), "point_folds_per_s": folds / cpu_s,
                           "cores": int(threads)}}


def learned_leg(n_curves: int, n_plain: int) -> dict:
    """The learned proposer + refinement on the learned workload (16,384 gapped curves,
    200,000 frequencies each) against the plain full-grid search. The learned path runs
    on every curve; the plain search is timed on the first ``n_plain`` curves (one
    captured launch) and extrapolated to the batch in sum(N x F); agreement is measured on
    those curves."""
    import torch

    from cuperiod_amd.batch import engine
    from cuperiod_amd.core.config import BLSSettings
    from cuperiod_amd.learned.proposer import DEFAULT_WEIGHTS, LearnedProposer
    from cuperiod_amd.workloads import learned_batch

    settings = BLSSettings(n_durations=6)
    s0 = time.perf_counter()
    wl = learned_batch(n_curves)
    gen_s = time.perf_counter() - s0
    n_plain = min(n_plain, n_curves)
    s0 = time.perf_counter()
    cs = engine.pack_curves(wl.curves)
    prop = LearnedProposer()
    prep = prop.prepare(cs, wl.grids, settings)
    prep_s = time.perf_counter() - s0
    log(f"learned leg: {n_curves} curves generated in {gen_s:.0f} s, prepared in "
        f"{prep_s:.0f} s ({prep.n_chunks} coarse chunks)")
    # plain on the first n_plain curves, replayed from a captured graph
    pb = engine.pack(wl.curves[:n_plain], wl.grids[:n_plain], settings)
    graph = engine.capture(pb)
    graph.launch()
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    graph.launch()
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - s0
    plain = engine.finish(pb, graph.out, graph.summary)
    graph.close()
    del graph, pb
    torch.cuda.empty_cache()
    pts = np.array([c[0].size for c in wl.curves], dtype=np.float64)
    plain_batch_s = plain_s * pts.sum() / pts[:n_plain].sum()
    prop.search(prep, settings)  # warm-up
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    r = prop.search(prep, settings)
    torch.cuda.synchronize()
    learned_s = time.perf_counter() - s0
    got = r.best_period[:n_plain]
    identical = got == plain.best_period
    same_power = r.best_power[:n_plain] == plain.best_power
    keep = ~r.fallback
    return {
        "workload": f"learned batch: {n_curves} gapped curves (N ~14,000 over 1,000 d, "
                    "30% in seasonal gaps, a 0.2 d transit of depth 1.5e-3 at a period "
                    "log-uniform in [2, 20] d), 200,000 frequencies in [1/1000, 2] /d, 6 "
                    "segment durations",
        "value": round(n_curves / learned_s, 3),
        "unit": "light curves/s",
        "learned_s": round(learned_s, 3),
        "plain_value": round(n_curves / plain_batch_s, 3),
        "plain_timed_curves": n_plain,
        "plain_s_timed": round(plain_s, 3),
        "plain_batch_s_extrapolated": round(plain_batch_s, 2),
        "speedup_vs_plain": round(plain_batch_s / learned_s, 3),
        "best_period_match": float(np.mean(identical)),
        "best_period_match_sample": f"the first {n_plain} curves (plain search's best "
                                    "period, bit for bit)",
        "same_best_power_different_period": int(np.sum(same_power & ~identical)),
        "fallback_frac": float(np.mean(r.fallback)),
        "match_without_fallback": float(np.mean(identical[keep[:n_plain]]))
        if keep[:n_plain].any() else None,
        "fine_freqs_per_refined_curve": float(np.mean(r.n_fine[keep])) if keep.any() else None,
        "coarse": {"freqs": 4096, "chunks": prep.n_chunks,
                   "chunk_duration_d": prop.chunk_duration,
                   "bins_per_duration": prop.coarse_oversample},
        "windows": {"max": prop.max_windows, "mass": prop.window_mass,
                    "threshold": prop.threshold, "fine_per_window": prop.fine_per_window},
        "timings_s": {k: round(v, 4) for k, v in r.timings.items()},
        "host_prepare_s": round(prep_s, 1),
        "weights": DEFAULT_WEIGHTS.name,
        "training": _training_summary(DEFAULT_WEIGHTS),
    }


def _training_summary(weights) -> dict | None:
    report = weights.with_name(weights.stem + "_report.json")
    if not report.exists():
        return None
    with open(report) as fh:
        r = json.load(fh)
    last = r["history"][-1] if r.get("history") else {}
    per_curve = r.get("labelling_seconds", 0.0) / max(1, r.get("curves", 1))
    return {"curves": r.get("curves"), "epochs": r.get("epochs"),
            "val_kept_recall": last.get("kept_recall"),
            "val_kept_windows": last.get("mean_kept_windows"),
            "labelling_s": r.get("labelling_seconds"),
            "labelling_gpu_minutes_for_1M_curves": round(per_curve * 1e6 / 60.0, 1)}


def multiband_leg(n_stars: int, cpu: bool) -> dict:
    """Joint-ephemeris multi-band BLS: stars/s over one batched launch, the sweep's
    box-band evaluations, and the CPU oracle on two stars."""
    from cuperiod_amd import _lib
    from cuperiod_amd.core.config import BLSSettings
    from cuperiod_amd.core.lightcurve import LightCurve
    from cuperiod_amd.methods import bls as bls_mod
    from cuperiod_amd.methods._bls_core import _search_spec
    from cuperiod_amd.methods._bls_joint import joint_bands, joint_search_stars
    from cuperiod_amd.workloads import multiband_survey

    stars, truth = multiband_survey(n_stars)
    settings = BLSSettings(mb_ephemeris="shared")
    prepared = [joint_bands(*zip(*s)) for s in stars]
    t_all = np.sort(np.concatenate([b[0] for b in stars[0]]))
    grid = bls_mod.BLSMethod().default_grid(
        LightCurve.from_arrays(t_all, np.ones(t_all.size), domain="flux"), settings)
    segs = bls_mod._grid_segments(grid, settings)
    os_ = settings.bins_per_duration
    boxes = 0.0
    for periods, durations in segs:
        spec = _search_spec(periods, durations, os_)
        nb = np.minimum(np.ceil(periods / spec.bin_duration).astype(np.int64) + os_,
                        spec.width - 1)
        for d in spec.dur_bins:
            boxes += float(np.sum(np.where(nb > d, nb - d + 1, 0)))
    box_band = boxes * float(sum(p.n_bands for p in prepared))
    folds = float(grid.values.size) * float(sum(p.tau.size for p in prepared))
    ctx = _lib.context(0)
    ctx.lib.cpb_set_timing(ctx.handle, 1)
    joint_search_stars(prepared[:2], [segs] * 2, os_, objective="snr", backend="hip")
    s0 = time.perf_counter()
    found = joint_search_stars(prepared, [segs] * len(prepared), os_, objective="snr",
                               backend="hip")
    wall = time.perf_counter() - s0
    k_ms = _kernel_ms(ctx)
    periods = np.concatenate([p for p, _ in segs])
    hits = 0
    for f, p_true in zip(found, truth):
        if p_true is None:
            continue
        power = np.concatenate([r["power"] for r in f])
        p_best = periods[int(np.argmax(power))]
        hits += any(abs(p_best * k / p_true - 1.0) < 0.01 for k in (0.5, 1.0, 2.0))
    out = {
        "workload": f"joint-ephemeris multi-band BLS, {len(stars)} stars x 6 bands x 150 "
                    f"points over 1095 d, {grid.values.size} default-grid periods per star",
        "kernel": "bls_joint_group_kernel",
        "value": round(len(stars) / (k_ms * 1e-3), 3),
        "unit": "stars/s",
        "kernel_ms": round(k_ms, 3),
        "wall_s": round(wall, 3),
        "box_band_evals_per_s": box_band / (k_ms * 1e-3),
        "point_folds_per_s": folds / (k_ms * 1e-3),
        "transits_recovered": f"{hits}/{sum(t is not None for t in truth)}",
        "cpu_baseline": None,
    }
    if cpu:
        from oracle import default_threads
        from oracle import multiband as omb

        n = 2
        s0 = time.perf_counter()
        refs = [omb.joint_search(omb.joint_bands(*zip(*s)), segs, os_,
                                 threads=default_threads()) for s in stars[:n]]
        el = time.perf_counter() - s0
        diff = 0.0
        for f, ref in zip(found, refs):
            got = np.concatenate([r["power"] for r in f])
            want = np.concatenate([r["power"] for r in ref])
            diff = max(diff, float(np.max(np.abs(got - want))
                                   / max(1e-300, float(np.max(np.abs(want))))))
        out["cpu_baseline"] = {
            "value": round(n / el, 4), "unit": "stars/s", "cores": default_threads(),
            "kind": "port", "sample": f"the first {n} stars ({el:.1f} s)",
            "gpu_over_cpu": round(out["value"] / (n / el), 1),
            "max_rel_power_diff_of_max": diff,
        }
    return out


def tls_leg(n_curves: int, cpu: bool) -> dict:
    """TLS matched filter: curves/s over one batched launch, point-folds and template
    taps, and the CPU oracle on a few curves."""
    from cuperiod_amd import _lib
    from cuperiod_amd.core.config import TLSSettings
    from cuperiod_amd.core.lightcurve import LightCurve
    from cuperiod_amd.methods import tls as tls_mod
    from cuperiod_amd.workloads import tls_survey

    curves, truth = tls_survey(n_curves)
    settings = TLSSettings()
    method = tls_mod.TLSMethod()
    periods = [method.default_grid(LightCurve.from_arrays(t, y, e, domain="flux"),
                                   settings).period for t, y, e in curves]
    widths = tls_mod._duration_bins(settings)
    n_bins = settings.n_phase_bins
    folds = float(sum(p.size * t.size for p, (t, _, _) in zip(periods, curves)))
    taps = float(sum(p.size for p in periods)) * n_bins * float(sum(widths))
    ctx = _lib.context(0)
    ctx.lib.cpb_set_timing(ctx.handle, 1)
    tls_mod.tls_power_batch(curves[:2], periods[:2], settings=settings)
    s0 = time.perf_counter()
    res = tls_mod.tls_power_batch(curves, periods, settings=settings)
    wall = time.perf_counter() - s0
    k_ms = _kernel_ms(ctx)
    hits = 0
    for r, p, p_true in zip(res, periods, truth):
        if p_true is not None:
            p_best = p[int(np.argmax(r["sde"]))]
            hits += any(abs(p_best * k / p_true - 1.0) < 0.01 for k in (0.5, 1.0, 2.0))
    # FP64 operations: 4 per point-fold (divide, floor, subtract, scale), 6 per template
    # tap (three multiply-adds)
    flops = 4.0 * folds + 6.0 * taps
    out = {
        "workload": f"TLS, {len(curves)} curves x 20000 points over 27 d, "
                    f"{periods[0].size} default-grid periods, {n_bins} phase bins, "
                    f"widths {widths}",
        "kernel": "tls_kernel",
        "value": round(len(curves) / (k_ms * 1e-3), 3),
        "unit": "light curves/s",
        "kernel_ms": round(k_ms, 3),
        "wall_s": round(wall, 3),
        "point_folds_per_s": folds / (k_ms * 1e-3),
        "template_taps_per_s": taps / (k_ms * 1e-3),
        "fp64_tflops": round(flops / (k_ms * 1e-3) / 1e12, 4),
        "fp64_frac_of_78.6": round(flops / (k_ms * 1e-3) / 78.6e12, 5),
        "transits_recovered": f"{hits}/{sum(t is not None for t in truth)}",
        "cpu_baseline": None,
    }
    if cpu:
        from oracle import default_threads
        from oracle import tls as otls

        n = min(64, len(curves))
        s0 = time.perf_counter()
        refs = [otls.tls_power(t, y, e, p, threads=default_threads())
                for (t, y, e), p in zip(curves[:n], periods[:n])]
        el = time.perf_counter() - s0
        diff = max(float(np.max(np.abs(r["sr"] - g["sr"]) / np.maximum(1e-300, np.abs(r["sr"]))))
                   for r, g in zip(refs, res[:n]))
        out["cpu_baseline"] = {
            "value": round(n / el, 4), "unit": "light curves/s", "cores": default_threads(),
            "kind": "port", "sample": f"the first {n} curves ({el:.1f} s)",
            "gpu_over_cpu": round(out["value"] / (n / el), 1),
            "max_rel_sr_diff": diff,
        }
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--curves", type=int, default=4096)
    ap.add_argument("--freqs", type=int, default=200_000)
    ap.add_argument("--cpu-curves", type=int, default=64,
                    help="curves of the batch searched by the oracle (parity and CPU rate)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of autotuning it")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each step directly instead of replaying a captured graph")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned proposer + refinement leg")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the multi-band and TLS legs")
    ap.add_argument("--no-e2e", action="store_true", help="skip the batch_periodograms leg")
    ap.add_argument("--no-long", action="store_true", help="skip the long-baseline leg")
    ap.add_argument("--long-curves", type=int, default=8)
    ap.add_argument("--learned-curves", type=int, default=16384)
    ap.add_argument("--learned-plain-curves", type=int, default=2048)
    ap.add_argument("--e2e-chunk", type=int, default=256,
                    help="batch_periodograms chunk size (its own default, 256)")
    ap.add_argument("--mb-stars", type=int, default=32)
    ap.add_argument("--tls-curves", type=int, default=256)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 256 MiB device copy (the HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the batch "
              "shards by curve with no exchange between GPUs)", file=sys.stderr)
        return 2

    import torch

    from cuperiod_amd import _lib
    from cuperiod_amd.batch import engine
    from cuperiod_amd.core.config import BLSSettings
    from cuperiod_amd.workloads import survey_batch

    log(f"building workload: {args.curves} curves x {args.freqs} frequencies")
    wl = survey_batch(args.curves, args.freqs)
    settings = BLSSettings(n_durations=wl.n_durations)
    ctx = _lib.context(0)
    if args.calibrate_copy:
        src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy of {COPY_BYTES >> 20} MiB done")
        del src, dst
    t0 = time.perf_counter()
    pb = engine.pack(wl.curves, wl.grids, settings)
    pack_s = time.perf_counter() - t0
    log(f"packed in {pack_s:.1f} s: {pb.n_segments} segments, {pb.n_out} periods, "
        f"{pb.total_points} points")
    folds, boxes = count_work(pb.plans, [c[0].size for c in wl.curves], pb.oversample)
    log(f"work per step: {folds:.3e} point-folds, {boxes:.3e} box evaluations")
    device = torch.device("cuda", ctx.device)
    out = torch.empty((_lib.N_OUTPUTS, pb.n_out), dtype=torch.float64, device=device)
    summary = torch.empty((len(pb.plans), 4), dtype=torch.float64, device=device)
    launch = {"graph": not args.no_graph, "shape": None}
    if not args.no_tune:
        from cuperiod_amd.batch import autotune

        sample = autotune.sample_batch(pb, settings, 64)
        tr = autotune.tune(pb, sample, use_cache=False)
        launch.update(shape=vars(tr.shape), tune_s=round(tr.seconds, 2),
                      tune_sample="first 64 curves, full grids",
                      tune_outputs_identical=tr.identical, candidates_ms=tr.timings_ms)
        log(f"autotuned launch shape {tr.shape} in {tr.seconds:.1f} s "
            f"(outputs identical across shapes: {tr.identical})")
        del sample
    ctx.lib.cpb_set_timing(ctx.handle, 1)
    if args.no_graph:
        def step():
            engine.run_packed(pb, out=out, summary=summary)
    else:
        graph = engine.capture(pb, out, summary)
        step = graph.launch
    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warm-up step {w + 1}/{args.warmup} done")
    kernel_ms = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s in range(args.steps):
        step()
        kernel_ms.append(_kernel_ms(ctx))
        log(f"step {s + 1}/{args.steps}: {'graph' if not args.no_graph else 'main kernel'} "
            f"{kernel_ms[-1]:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    res = engine.finish(pb, out, summary)
    ms_per_step = 1000.0 * elapsed / args.steps
    curves_per_s = len(pb.plans) / (elapsed / args.steps)
    k_ms = float(np.mean(kernel_ms))
    flops = 4.0 * folds + 8.0 * boxes
    achieved = flops / (k_ms * 1e-3)
    measured = MEASURED.get((len(pb.plans), args.freqs), {})
    issue = None
    if "valu_wave_instr" in measured:
        lane_rate = 64.0 * measured["valu_wave_instr"] / (k_ms * 1e-3)
        issue = {"valu_wave_instructions_per_step": measured["valu_wave_instr"],
                 "lane_instructions_per_s": lane_rate, "issue_limit": VALU_ISSUE,
                 "frac": round(lane_rate / VALU_ISSUE, 4),
                 "per_box_eval": round(64.0 * measured["valu_wave_instr"] / boxes, 2)}
    line = {
        "metric": METRIC,
        "value": round(curves_per_s, 3),
        "unit": "light curves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 phase, int64 fixed-point bins, f32 screen, f64 final",
        "data": "synthetic",
        "config": {
            "workload": f"plain batched BLS, {len(pb.plans)} curves x {args.freqs} "
                        "linspace frequencies in [1/baseline, 2] /d, 6 segment durations",
            "curves": len(pb.plans),
            "freqs_per_curve": args.freqs,
            "point_folds_per_step": folds,
            "box_evals_per_step": boxes,
            "host_pack_s": round(pack_s, 3),
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved / 1e12, 4),
            "peak": PEAK_VALU_FLOPS / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_VALU_FLOPS, 5),
            "traffic": measured.get("traffic"),
            "issue": issue,
        },
        "detail": {
            "kernel": "bls_main_kernel",
            "kernel_ms_mean": round(k_ms, 3),
            "point_folds_per_s": folds / (k_ms * 1e-3),
            "box_evals_per_s": boxes / (k_ms * 1e-3),
            "lds_bytes_per_s": 16.0 * (folds + boxes) / (k_ms * 1e-3),
            "lds_frac": 16.0 * (folds + boxes) / (k_ms * 1e-3) / PEAK_LDS,
            "best_period_sample": [float(x) for x in res.best_period[:4]],
        },
        "launch": launch,
        "parity": None,
        "cpu_baseline": None,
        "end_to_end": None,
        "long_baseline": None,
        "learned": None,
        "max_error_vs_fp64": None,
        "next_rows": None,
    }
    if not args.no_cpu and args.cpu_curves > 0:
        n = min(args.cpu_curves, len(pb.plans))
        log(f"oracle on {n} curves (parity and CPU rate) ...")
        parity, cpu = parity_and_cpu_leg(wl, settings, pb, n)
        cpu["extrapolated_batch_s"] = round(folds / cpu["sample_point_folds"] * n / cpu["value"], 1)
        cpu["gpu_over_cpu"] = round(curves_per_s / cpu["value"], 1)
        cpu["calibration"] = calibration_leg()
        line["parity"], line["cpu_baseline"] = parity, cpu
        log(f"parity: identical {parity['best_period_identical']:.4f}, ties "
            f"{parity['best_period_equal_point_ties']}, max rel err "
            f"{parity['max_rel_power_error']:.3g}; CPU {cpu['value']} curves/s")
    if not args.no_e2e:
        log("end-to-end batch_periodograms ...")
        line["end_to_end"] = end_to_end_leg(wl, settings, args.e2e_chunk)
        log(f"end-to-end: {line['end_to_end']['value']} curves/s")
    if not args.no_long:
        log("long-baseline batch ...")
        line["long_baseline"] = long_baseline_leg(args.long_curves, not args.no_cpu)
        log(f"long-baseline: {line['long_baseline']['value']} curves/s")
    if not args.no_cpu:
        log("FP64 error per workload ...")
        line["max_error_vs_fp64"] = fp64_error_leg(line["parity"], line["long_baseline"])
        log(f"FP64 errors: {json.dumps(line['max_error_vs_fp64'])}")
    if not args.no_learned:
        from cuperiod_amd.learned.proposer import DEFAULT_WEIGHTS

        if DEFAULT_WEIGHTS.exists():
            log("learned proposer + refinement ...")
            line["learned"] = learned_leg(args.learned_curves, args.learned_plain_curves)
            lr = line["learned"]
            log(f"learned: {lr['value']} curves/s, {lr['speedup_vs_plain']}x plain, match "
                f"{lr['best_period_match']:.4f}, fallback {lr['fallback_frac']:.4f}")
    if not args.no_next:
        log("multi-band joint search ...")
        mb = multiband_leg(args.mb_stars, not args.no_cpu)
        log(f"multi-band: {mb['value']} stars/s ({mb['kernel_ms']} ms)")
        log("TLS ...")
        tl = tls_leg(args.tls_curves, not args.no_cpu)
        log(f"TLS: {tl['value']} curves/s ({tl['kernel_ms']} ms)")
        line["next_rows"] = {"multiband": mb, "tls": tl}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
