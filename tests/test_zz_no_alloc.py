"""The allocation-free batched search (``cpb_bls_search_batch_static``) against the
allocating one (``cpb_bls_search_batch``), through the C ABI.

Both run the same kernels on the same packed inputs, so every output must be bit for
bit the same (the engine's results are integer-exact sums rounded once, and the peak
rounding is deterministic). Two calls in a row, the second on a smaller batch, check
that the static scratch the two calls share carries nothing from one call into the next;
a search beyond the scratch's size is refused with the resource error.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cuperiod_amd import _lib
from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.core.errors import BackendUnavailableError
from cuperiod_amd.workloads import survey_curves

pytestmark = pytest.mark.gpu


def _vp(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _search(entry: str, pb: engine.PackedBatch) -> np.ndarray:
    import torch

    ctx = _lib.context(0)
    device = torch.device("cuda", ctx.device)
    out = torch.full((_lib.N_OUTPUTS, max(1, pb.n_out)), float("nan"), dtype=torch.float64,
                     device=device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    with ctx.lock:
        st = getattr(ctx.lib, entry)(
            ctx.handle, stream, _vp(pb.tau), _vp(pb.yw), _vp(pb.ivar), pb.total_points,
            pb.curves, len(pb.plans), _vp(pb.periods), _vp(pb.dur_bins), pb.segments,
            pb.n_segments, pb.oversample, pb.obj_flag, _vp(out), pb.n_out)
        ctx.check(st, entry)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("objective", ["snr", "likelihood"])
def test_static_search_equals_the_allocating_search(objective):
    settings = BLSSettings(n_durations=6, objective=objective)
    # the second call is smaller: the scratch holds the first call's larger tables
    for n_curves, seed, n_freq in ((24, 11, 20_000), (5, 12, 8_000)):
        print(f"case static scratch: {n_curves} curves x {n_freq} frequencies, "
              f"objective {objective}", flush=True)
        curves, grids, _ = survey_curves(n_curves, seed=seed, n_freq=n_freq)
        pb = engine.pack(curves, grids, settings)
        static = _search("cpb_bls_search_batch_static", pb)
        alloc = _search("cpb_bls_search_batch", pb)
        assert not np.isnan(static).any()
        np.testing.assert_array_equal(static, alloc)


def test_static_search_refuses_more_than_its_scratch():
    # 5,000,000 points need 80 MB (76 MiB) of fixed-point copies alone, beyond the
    # 64 MiB array
    print("case static scratch: 5,000,000 points (refused)", flush=True)
    rng = np.random.default_rng(3)
    n = 5_000_000
    t = np.sort(rng.uniform(0.0, 100.0, n))
    y = 1.0 + 1e-3 * rng.standard_normal(n)
    curves = [(t, y, np.full(n, 1e-3))]
    grid = engine.LinspaceGrid(0.05, 1.0, 16)
    pb = engine.pack(curves, [grid], BLSSettings(n_durations=2))
    with pytest.raises(BackendUnavailableError, match="static scratch"):
        _search("cpb_bls_search_batch_static", pb)
    # the allocating entry point takes it
    out = _search("cpb_bls_search_batch", pb)
    assert np.isfinite(out[0]).all()
