"""Launch shapes and HIP-graph replays of the batched search give the same outputs.

Every shape the autotuner may pick (workgroup size, periods per work item, persistent
grid) must reproduce the default shape's outputs bit for bit; a captured graph is
replayed several times, with other searches and synchronisations in between, and every
replay must equal the eager search.
"""

from __future__ import annotations

import numpy as np
import pytest

from cuperiod_amd.batch import autotune, engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.workloads import survey_curves

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def packed():
    curves, grids, _ = survey_curves(8, seed=4242, n_freq=20_000)
    settings = BLSSettings(n_durations=6)
    return settings, engine.pack(curves, grids, settings), curves


def test_launch_shapes_are_bitwise_identical(packed):
    import torch

    settings, pb, _ = packed
    autotune.apply(autotune.DEFAULT_SHAPE)
    ref_out, ref_sum = engine.run_packed(pb)
    ref_out, ref_sum = ref_out.clone(), ref_sum.clone()
    try:
        for shape in autotune.CANDIDATES:
            autotune.apply(shape)
            out, summ = engine.run_packed(pb)
            torch.cuda.synchronize()
            assert torch.equal(out, ref_out), shape
            assert torch.equal(summ, ref_sum), shape
    finally:
        autotune.apply(autotune.DEFAULT_SHAPE)


def test_graph_replays_equal_the_eager_search(packed):
    import torch

    settings, pb, curves = packed
    eager_out, eager_sum = engine.run_packed(pb)
    eager_out, eager_sum = eager_out.clone(), eager_sum.clone()
    graph = engine.capture(pb)
    other = engine.pack(curves[:3], [engine.LinspaceGrid(0.05, 1.0, 5000)] * 3, settings)
    try:
        for _ in range(3):
            graph.out.fill_(np.nan)
            graph.summary.zero_()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(graph.out, eager_out)
            assert torch.equal(graph.summary, eager_sum)
            engine.run_packed(other)  # eager work of the program between replays
            torch.cuda.synchronize()
    finally:
        graph.close()


def test_autotuner_picks_a_candidate(packed, tmp_path, monkeypatch):
    settings, pb, _ = packed
    monkeypatch.setenv("CUPERIOD_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    sample = autotune.sample_batch(pb, settings, 4)
    shapes = autotune.CANDIDATES[:4]
    try:
        res = autotune.tune(pb, sample, candidates=shapes, repeats=1)
        assert res.shape in shapes and res.identical and not res.cached
        again = autotune.tune(pb, sample, candidates=shapes, repeats=1)
        assert again.cached and again.shape == res.shape
    finally:
        autotune.apply(autotune.DEFAULT_SHAPE)


def test_plateau_choice_does_not_depend_on_the_launch_shape():
    # One deep point alone in the best box at thousands of periods: more periods tie at
    # the maximum than the reference rounding searches again, and which of them it takes
    # must be the same for every launch shape and every run.
    import torch

    from cuperiod_amd import _lib

    rng = np.random.default_rng(11)
    t = np.sort(rng.uniform(0.0, 40.0, 60))
    y = 1.0 + 1e-4 * rng.standard_normal(60)
    y[17] -= 0.05
    settings = BLSSettings(n_durations=6)
    pb = engine.pack([(t, y, np.full(60, 1e-4))], [engine.LinspaceGrid(1 / 40, 2.0, 20_000)],
                     settings)
    ctx = _lib.context(0)
    with ctx.peak_rounding(False):
        raw, _ = engine.run_packed(pb)
    power = raw[0].cpu().numpy()
    assert int((power == np.nanmax(power)).sum()) > 64
    autotune.apply(autotune.DEFAULT_SHAPE)
    ref_out, ref_sum = engine.run_packed(pb)
    ref_out, ref_sum = ref_out.clone(), ref_sum.clone()
    try:
        for shape in autotune.CANDIDATES:
            autotune.apply(shape)
            for _ in range(2):
                out, summ = engine.run_packed(pb)
                torch.cuda.synchronize()
                assert torch.equal(out, ref_out), shape
                assert torch.equal(summ, ref_sum), shape
    finally:
        autotune.apply(autotune.DEFAULT_SHAPE)
