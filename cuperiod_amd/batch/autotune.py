"""Launch-shape autotuner for the batched BLS engine.

The engine's results do not depend on its launch shape (``cpb_set_launch``): the
workgroup size (128, 256 or 512 threads), the periods per work item (the tile) and a
persistent grid of a few workgroups per compute unit against one workgroup per item.
Their speed does, through occupancy, barrier cost and how evenly uneven curves spread
over the 256 compute units. The tuner times every candidate shape on a sample of the
batch's curves (their full grids), keeps the fastest, and caches it per device and
workload signature in a JSON file (``$CUPERIOD_AMD_TUNE_CACHE``, default
``~/.cache/cuperiod_amd/autotune.json``).
"""

from __future__ import annotations

import ctypes
import json
import os
import time
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

from cuperiod_amd import _lib
from cuperiod_amd.batch import engine


@dataclass(frozen=True)
class LaunchShape:
    """``threads`` per workgroup, ``tile`` periods per work item (0: automatic),
    ``grid_per_cu`` persistent workgroups per CU (0: one workgroup per item)."""

    threads: int = 256
    tile: int = 0
    grid_per_cu: int = 0


DEFAULT_SHAPE = LaunchShape()

#: Shapes the tuner tries.
CANDIDATES: tuple[LaunchShape, ...] = tuple(
    LaunchShape(t, tile, g) for t in (128, 256, 512) for tile in (0, 8, 32) for g in (0, 8)
)


def apply(shape: LaunchShape) -> None:
    """Make ``shape`` the launch shape of later searches on device 0's context."""
    ctx = _lib.context(0)
    with ctx.lock:
        ctx.check(ctx.lib.cpb_set_launch(ctx.handle, shape.threads, shape.tile,
                                         shape.grid_per_cu), "launch shape")


def _bucket(x: float) -> int:
    return int(2 ** round(np.log2(max(1.0, x))))


def signature(pb: engine.PackedBatch) -> str:
    """Workload key: device, curve-size and phase-span buckets, durations, objective and
    accumulation, and the periods per curve bucket."""
    ctx = _lib.context(0)
    import torch

    props = torch.cuda.get_device_properties(ctx.device)
    npts = np.median(pb.curve_set.n_points) if pb.curve_set is not None else 0.0
    spans = [sg.p_max / (float(sg.durations.min()) / pb.oversample)
             for plan in pb.plans[:256] for sg in plan.segments]
    ndur = max((len(sg.durations) for plan in pb.plans[:256] for sg in plan.segments),
               default=0)
    per = np.median(pb.curve_count) if len(pb.curve_count) else 0.0
    arch = getattr(props, "gcnArchName", props.name).split(":")[0]
    return (f"{arch}/cu{props.multi_processor_count}/pts{_bucket(npts)}"
            f"/span{_bucket(float(np.median(spans)) if spans else 0.0)}/dur{ndur}"
            f"/per{_bucket(per)}/obj{pb.obj_flag}/acc{pb.accumulation}")


def cache_path() -> Path:
    env = os.environ.get("CUPERIOD_AMD_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "cuperiod_amd" / "autotune.json"


def _load_cache(path: Path) -> dict:
    try:
        data = json.loads(path.read_text())
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: Path, key: str, shape: LaunchShape, timings: dict) -> None:
    data = _load_cache(path)
    data[key] = {"shape": asdict(shape), "ms": timings}
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_name(path.name + f".{os.getpid()}.tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def _time_shape(pb: engine.PackedBatch, shape: LaunchShape, out, summary, repeats: int) -> float:
    import torch

    ctx = _lib.context(0)
    apply(shape)
    engine.run_packed(pb, out=out, summary=summary)  # warm-up (and first-use setup)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(repeats):
        engine.run_packed(pb, out=out, summary=summary)
        ms = ctypes.c_double()
        ctx.check(ctx.lib.cpb_main_kernel_ms(ctx.handle, ctypes.byref(ms)), "timing")
        best = min(best, ms.value)
    return best


@dataclass
class TuneResult:
    shape: LaunchShape
    signature: str
    timings_ms: dict
    seconds: float
    cached: bool
    identical: bool


def tune(pb: engine.PackedBatch, sample: engine.PackedBatch, *,
         candidates: Sequence[LaunchShape] = CANDIDATES, repeats: int = 2,
         use_cache: bool = True) -> TuneResult:
    """The fastest shape for ``pb``'s workload, timed on ``sample`` (a few of its curves
    packed with their full plans); applied before returning. Every candidate's outputs
    are compared with the default shape's (``identical``)."""
    import torch

    key = signature(pb)
    path = cache_path()
    if use_cache:
        hit = _load_cache(path).get(key)
        if hit and isinstance(hit.get("shape"), dict):
            shape = LaunchShape(**hit["shape"])
            apply(shape)
            return TuneResult(shape, key, hit.get("ms", {}), 0.0, True, True)
    ctx = _lib.context(0)
    t0 = time.perf_counter()
    device = sample.tau.device
    out = torch.empty((_lib.N_OUTPUTS, max(1, sample.n_out)), dtype=torch.float64, device=device)
    summary = torch.empty((len(sample.plans), 4), dtype=torch.float64, device=device)
    with ctx.lock:
        ctx.check(ctx.lib.cpb_set_timing(ctx.handle, 1), "timing")
    apply(DEFAULT_SHAPE)
    engine.run_packed(sample, out=out, summary=summary)
    ref = out.clone()
    identical = True
    timings: dict[str, float] = {}
    best_shape, best_ms = DEFAULT_SHAPE, float("inf")
    for shape in candidates:
        ms = _time_shape(sample, shape, out, summary, repeats)
        timings[f"{shape.threads}/{shape.tile}/{shape.grid_per_cu}"] = round(ms, 3)
        identical &= bool(torch.equal(out, ref))
        if ms < best_ms:
            best_shape, best_ms = shape, ms
    apply(best_shape)
    if use_cache:
        _store(path, key, best_shape, timings)
    return TuneResult(best_shape, key, timings, time.perf_counter() - t0, False, identical)


def sample_batch(pb: engine.PackedBatch, settings, n: int = 64) -> engine.PackedBatch:
    """The first ``n`` curves of ``pb`` with their full plans, for timing."""
    from cuperiod_amd.batch.engine import CurveSet

    cs = pb.curve_set
    n = min(n, len(pb.plans))
    last = int(cs.point_base[n - 1] + cs.n_points[n - 1])
    sub = CurveSet(
        tau=cs.tau[:last], yw=cs.yw[:last], ivar=cs.ivar[:last], total_points=last,
        curves=cs.curves, curves32=cs.curves32, t_min=cs.t_min[:n], baseline=cs.baseline[:n],
        n_points=cs.n_points[:n], point_base=cs.point_base[:n], times=cs.times[:n],
    )
    return engine.pack_plans(sub, pb.plans[:n], settings)


__all__ = ["CANDIDATES", "DEFAULT_SHAPE", "LaunchShape", "TuneResult", "apply", "cache_path",
           "sample_batch", "signature", "tune"]
