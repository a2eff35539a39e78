"""Benchmark: batched plain BLS on one MI355X (and the CPU oracle beside it).

Workload (default): the 4,096-curve survey batch -- N log-uniform in [100, 5,000],
baselines 50-2,000 d, 200,000 frequencies per curve in [1/baseline, 2] /d, 6
period-scaled durations per log segment (``cuperiod_amd.workloads.survey_batch``).

A step is one batched search of every curve (fold, box scan, argmax, per-curve
summary with the best period) with the inputs already resident in HBM. W untimed warm-up
steps, then K timed steps between two device synchronisations. One JSON line is printed.

``roofline``: the main kernel's algorithmic FLOP (4 FP64 per point-fold, 8 FP32 per box
evaluation) per launch over its mean HIP-event duration, against the 157.3 TFLOP/s FP32
peak; ``detail`` adds point-folds/s, box-evaluations/s and the LDS bytes they imply
(16 B each) against 256 CUs x 128 B/clk x 2.4 GHz.

``cpu_baseline``: the FP64 C++ oracle (the reference's numba kernel restated, with its
screen) on the first ``--cpu-curves`` curves of the same batch, all host threads.

``next_rows``: the joint-ephemeris multi-band search (``multiband_survey``: 6 bands x 150
points over 1,095 d per star, the default BLS grid) and the TLS matched filter
(``tls_survey``: 20,000 points over 27 d per curve, the default TLS grid), each timed over
one batched launch and beside its own CPU oracle.
"""

from __future__ import annotations

import argparse
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
PEAK_FP32 = 157.3e12
PEAK_LDS = 256 * 128 * 2.4e9
PEAK_HBM = 8.0e12
COPY_BYTES = 256 << 20
#: HBM bytes per step of bls_main_kernel (its launch classes), by (curves, frequencies
#: per curve) of the survey batch: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes of
#: this script (--steps 1 --warmup 0 --no-cpu --no-learned --calibrate-copy), each counter
#: scaled by what it reads for the plain 256 MiB copy (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000).
MEASURED_TRAFFIC = {(4096, 200_000): 7.539e9 + 58.69e9}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def count_work(plans, n_points, oversample) -> tuple[float, float]:
    """Exact point-folds and box evaluations of one launch over a batch."""
    from cuperiod_amd.methods._bls_core import _duration_bins

    folds = 0.0
    boxes = 0.0
    for plan, n in zip(plans, n_points):
        for sg in plan.segments:
            if sg.mode == 0:
                periods = sg.periods
            else:
                k = np.arange(sg.n_periods, dtype=np.int64)
                i = sg.lin_num - 1 - (sg.lin_k0 + k)
                f = i.astype(np.float64) * sg.lin_step + sg.lin_start
                if sg.mode == 1:
                    f = np.where(i == sg.lin_num - 1, sg.lin_stop, f)
                else:
                    f = np.where(i == 0, sg.lin_start, f)
                periods = 1.0 / f
            bd = float(sg.durations.min()) / oversample
            width = int(np.ceil(sg.p_max / bd)) + oversample + 1
            nb = np.minimum(np.ceil(periods / bd).astype(np.int64) + oversample, width - 1)
            folds += float(n) * sg.n_periods
            for d in _duration_bins(sg.durations, bd):
                m = nb[nb > d]
                boxes += float((m - d + 1).sum())
    return folds, boxes


def cpu_baseline(wl, settings, n_curves: int):
    """The oracle on the first ``n_curves`` curves (same grids, same settings)."""
    import oracle
    from cuperiod_amd.methods.bls import _grid_segments

    threads = oracle.default_threads()
    jobs = []
    for c in range(n_curves):
        t, y, dy = wl.curves[c]
        segs = _grid_segments(wl.grids[c].as_gridspec(), settings)
        jobs.append((t, y, dy, segs))
    oracle.bls_power(jobs[0][0][:50], jobs[0][1][:50], jobs[0][2][:50],
                     np.array([1.0, 2.0]), np.array([0.1]), 10)  # load + warm
    results = []
    t0 = time.perf_counter()
    for t, y, dy, segs in jobs:
        parts = [oracle.bls_power(t, y, dy, p, d, settings.bins_per_duration,
                                  objective=settings.objective, threads=threads)
                 for p, d in segs]
        results.append(parts)
    elapsed = time.perf_counter() - t0
    best, powers = [], []
    from cuperiod_amd.core.result import Periodogram

    for parts in results:
        power = np.nan_to_num(np.concatenate([p["power"] for p in parts]), nan=0.0,
                              posinf=0.0, neginf=0.0)
        period = np.concatenate([p["period"] for p in parts])
        pg = Periodogram.from_spectrum(method="BLS", backend="oracle", frequency=1.0 / period,
                                       power=power, objective_sense="max", n_samples=1,
                                       baseline=1.0)
        best.append(pg.best_period())
        powers.append(power)
    return elapsed, threads, np.asarray(best), powers


def learned_leg(pb, wl, settings, plain, plain_s: float, steps: int) -> dict:
    """The learned proposer + refinement on the same batch, against the plain search:
    speed, and agreement of the best period (identical, or the same best power, i.e. an
    exact tie the plain search resolved to another sample)."""
    import torch

    from cuperiod_amd.learned.proposer import DEFAULT_WEIGHTS, LearnedProposer

    prop = LearnedProposer()
    t0 = time.perf_counter()
    prep = prop.prepare(pb.curve_set, wl.grids, settings, full=pb.plans)
    prep_s = time.perf_counter() - t0
    prop.search(prep, settings)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = prop.search(prep, settings)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    identical = r.best_period == plain.best_period
    same_power = r.best_power == plain.best_power
    keep = ~r.fallback
    return {
        "value": round(len(pb.plans) / el, 3),
        "unit": "light curves/s",
        "ms_per_step": round(1000.0 * el, 3),
        "speedup_vs_plain": round(plain_s / el, 3),
        "best_period_match": float(np.mean(identical | same_power)),
        "best_period_identical": float(np.mean(identical)),
        "fallback_frac": float(np.mean(r.fallback)),
        "match_without_fallback": float(np.mean((identical | same_power)[keep])) if keep.any() else None,
        "fine_freqs_per_refined_curve": float(np.mean(r.n_fine[keep])) if keep.any() else None,
        "coarse_freqs": 4096,
        "windows": {"max": prop.max_windows, "mass": prop.window_mass,
                    "threshold": prop.threshold, "fine_per_window": prop.fine_per_window},
        "timings_s": {k: round(v, 4) for k, v in r.timings.items()},
        "prepare_s": round(prep_s, 3),
        "weights": DEFAULT_WEIGHTS.name,
    }


def _kernel_ms(ctx) -> float:
    import ctypes

    ms = ctypes.c_double()
    ctx.check(ctx.lib.cpb_main_kernel_ms(ctx.handle, ctypes.byref(ms)), "timing")
    return ms.value


def multiband_leg(n_stars: int, cpu: bool) -> dict:
    """Joint-ephemeris multi-band BLS: stars/s over one batched launch, the sweep's
    box-band evaluations, and the CPU oracle on one star."""
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
    k_bands = [p.n_bands for p in prepared]
    box_band = boxes * float(sum(k_bands))
    folds = float(grid.values.size) * float(sum(p.tau.size for p in prepared))
    ctx = _lib.context(0)
    ctx.lib.cpb_set_timing(ctx.handle, 1)
    joint_search_stars(prepared[:2], [segs] * 2, os_, objective="snr", backend="hip")
    t0 = time.perf_counter()
    found = joint_search_stars(prepared, [segs] * len(prepared), os_, objective="snr",
                               backend="hip")
    wall = time.perf_counter() - t0
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
        "kernel": "bls_joint_kernel",
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
        from oracle import multiband as omb
        from oracle import default_threads

        n = 2
        t0 = time.perf_counter()
        refs = [omb.joint_search(omb.joint_bands(*zip(*s)), segs, os_,
                                 threads=default_threads()) for s in stars[:n]]
        el = time.perf_counter() - t0
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
    t0 = time.perf_counter()
    res = tls_mod.tls_power_batch(curves, periods, settings=settings)
    wall = time.perf_counter() - t0
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
        t0 = time.perf_counter()
        refs = [otls.tls_power(t, y, e, p, threads=default_threads())
                for (t, y, e), p in zip(curves[:n], periods[:n])]
        el = time.perf_counter() - t0
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
    ap.add_argument("--cpu-curves", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of autotuning it")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch each step directly instead of replaying a captured graph")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned proposer + refinement leg")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the multi-band and TLS legs")
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
    log(f"packed in {time.perf_counter() - t0:.1f} s: {pb.n_segments} segments, "
        f"{pb.n_out} periods, {pb.total_points} points")
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
        import ctypes

        ms = ctypes.c_double()
        ctx.check(ctx.lib.cpb_main_kernel_ms(ctx.handle, ctypes.byref(ms)), "timing")
        kernel_ms.append(ms.value)
        log(f"step {s + 1}/{args.steps}: {'graph' if not args.no_graph else 'main kernel'} "
            f"{ms.value:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    res = engine.finish(pb, out, summary)
    ms_per_step = 1000.0 * elapsed / args.steps
    curves_per_s = len(pb.plans) / (elapsed / args.steps)
    k_ms = float(np.mean(kernel_ms))
    flops = 4.0 * folds + 8.0 * boxes
    achieved = flops / (k_ms * 1e-3)
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
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved / 1e12, 4),
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32, 5),
            "traffic": MEASURED_TRAFFIC.get((len(pb.plans), args.freqs)),
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
        "learned": None,
        "next_rows": None,
        "cpu_baseline": None,
    }
    from cuperiod_amd.learned.proposer import DEFAULT_WEIGHTS

    if not args.no_learned and DEFAULT_WEIGHTS.exists():
        log("learned proposer + refinement ...")
        line["learned"] = learned_leg(pb, wl, settings, res, elapsed / args.steps,
                                      max(1, min(args.steps, 2)))
        log(f"learned: {line['learned']['value']} curves/s, match "
            f"{line['learned']['best_period_match']:.4f}")
    if not args.no_next:
        log("multi-band joint search ...")
        mb = multiband_leg(args.mb_stars, not args.no_cpu)
        log(f"multi-band: {mb['value']} stars/s ({mb['kernel_ms']} ms)")
        log("TLS ...")
        tl = tls_leg(args.tls_curves, not args.no_cpu)
        log(f"TLS: {tl['value']} curves/s ({tl['kernel_ms']} ms)")
        line["next_rows"] = {"multiband": mb, "tls": tl}
    if not args.no_cpu and args.cpu_curves > 0:
        n = min(args.cpu_curves, len(pb.plans))
        log(f"CPU oracle on {n} curves ...")
        el, threads, best_cpu, cpu_powers = cpu_baseline(wl, settings, n)
        identical = best_cpu == res.best_period[:n]
        # A different sample of a flat-topped peak is an exact tie: the oracle's power
        # at the GPU's choice equals the oracle's maximum to 1e-12.
        tie = np.array([abs(p[int(res.best_index[c])] - p.max()) <= 1e-12 * abs(p.max())
                        for c, p in enumerate(cpu_powers)])
        match = float(np.mean(identical | tie))
        line["cpu_baseline"] = {
            "value": round(n / el, 5),
            "unit": "light curves/s",
            "cores": int(threads),
            "kind": "port",
            "sample": f"first {n} curves of the same batch, same grids and settings "
                      f"({el:.1f} s)",
            "best_period_match": match,
            "best_period_identical": float(np.mean(identical)),
            "gpu_over_cpu": round(curves_per_s / (n / el), 1),
        }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
