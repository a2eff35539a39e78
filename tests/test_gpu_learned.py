"""The learned path's refinement and fallback against the plain full-grid search.

Refinement searches a subset of the plain grid with the plain search's segment binning,
so its outputs at those periods are the plain outputs bit for bit; a curve that falls
back is the plain search. These properties hold for any network weights.
"""

from __future__ import annotations

import numpy as np
import pytest

from cuperiod_amd import _lib
from cuperiod_amd.batch import engine
from cuperiod_amd.core.config import BLSSettings
from cuperiod_amd.learned.network import init_parameters
from cuperiod_amd.learned.proposer import LearnedProposer
from cuperiod_amd.workloads import survey_curves

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    curves, grids, _ = survey_curves(10, seed=991, n_freq=30_000)
    settings = BLSSettings(n_durations=6)
    cs = engine.pack_curves(curves)
    prop = LearnedProposer(init_parameters(0), threshold=0.0, window_mass=1.0)
    prep = prop.prepare(cs, grids, settings)
    pb = engine.pack_plans(cs, prep.full, settings)
    plain = engine.finish(pb, *engine.run_packed(pb), spectra=True)
    return settings, cs, prop, prep, plain


def test_window_plan_outputs_equal_the_full_search(batch):
    # A period's exact-sum outputs do not depend on what else is searched with it.
    # (With the reference rounding of each segment's near-maximum periods, the periods
    # of a window that holds its segment's best are rounded as in the full search.)
    settings, cs, prop, prep, _ = batch
    ranges = [(0, 700), (12_000, 13_500), (29_000, 30_000)]
    plans = [engine.window_plan(prep.full[c], ranges) for c in range(len(prep.full))]
    ctx = _lib.context(0)
    with ctx.peak_rounding(False):
        pb = engine.pack_plans(cs, prep.full, settings)
        plain = engine.finish(pb, *engine.run_packed(pb), spectra=True)
        pb = engine.pack_plans(cs, plans, settings)
        res = engine.finish(pb, *engine.run_packed(pb), spectra=True)
    idx = np.concatenate([np.arange(a, b) for a, b in ranges])
    for c in range(len(plans)):
        for name, row in res.spectra[c].items():
            np.testing.assert_array_equal(row, plain.spectra[c][name][idx], err_msg=name)
        np.testing.assert_array_equal(plans[c].periods(), prep.full[c].periods()[idx])


def test_refinement_never_beats_the_plain_maximum(batch):
    settings, cs, prop, prep, plain = batch
    r = prop.search(prep, settings)
    assert not r.fallback.any()
    assert np.all(r.n_fine <= 32 * 2001)
    assert np.all(r.best_power <= plain.best_power)
    same = r.best_power == plain.best_power
    for c in np.flatnonzero(same):
        k = int(np.argmin(np.abs(prep.full[c].periods() - r.best_period[c])))
        p = np.nan_to_num(plain.spectra[c]["power"], nan=0.0, posinf=0.0, neginf=0.0)
        assert p[k] == p.max()


def test_full_fallback_is_the_plain_search(batch):
    settings, cs, prop, prep, plain = batch
    strict = LearnedProposer(init_parameters(0), threshold=2.0)
    r = strict.search(prep, settings)
    assert r.fallback.all()
    np.testing.assert_array_equal(r.best_period, plain.best_period)
    np.testing.assert_array_equal(r.best_power, plain.best_power)
