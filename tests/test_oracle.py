"""Pins the FP64 oracle before anything is checked against it.

Each test checks the oracle against something it does not share code with: hand
arithmetic, exactly summed brute force, the reference's own known answers and
invariances, or closed-form statistics. A plausible slip in the restatement (a dropped
``+1`` bin offset, the wrap copy done as an add, a swapped in/out sum, a wrong
tie-break) fails at least one of them.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle.bruteforce import _members, bls_bruteforce, box_table
from synth import adversarial_cases, synthetic_eclipser, two_runs


def test_duration_bins_known_answer():
    # tests/test_bls.py:213-215 of the reference: C round(), half away from zero.
    assert oracle.duration_bins(np.array([0.2, 0.25, 0.45]), 0.1) == [2, 3, 5]
    # Deduplicated in first-seen order, widths < 1 dropped.
    assert oracle.duration_bins(np.array([0.3, 0.1, 0.31, 0.04]), 0.1) == [3, 1]


def test_wrap_copy_hand_case():
    # P = 1.0, bd = 0.3, oversample = 2: n_bins = ceil(1/0.3) + 2 = 6; data bins 1..4;
    # the wrap COPIES bin 1 -> bin 4 and bin 2 -> bin 5, so the point in bin 4 is lost.
    tau = np.array([0.05, 0.35, 0.65, 0.95])
    bins, n_bins = _members(tau, 1.0, 0.3, 2, 8)
    assert n_bins == 6
    assert bins[0] == [] and bins[1] == [0] and bins[2] == [1] and bins[3] == [2]
    assert bins[4] == [0] and bins[5] == [1] and bins[6] == []


def test_wrap_copy_changes_the_oracle_result():
    # A dip whose only point sits in the overwritten last data bin is invisible;
    # moving that point into bin 1 makes it the deepest box. An add instead of a copy,
    # or a missing wrap, fails one of the two.
    t = np.array([0.05, 0.35, 0.65, 0.95, 1.05, 1.35, 1.65, 2.0])
    y = np.array([0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 0.0])
    ivar = np.ones_like(t)
    common = dict(bin_duration=0.3, dur_bins=[1], oversample=2, width=8, obj_flag=1)
    ref = oracle.bls_search(t, y, ivar, np.array([1.0]), **common)
    bf = bls_bruteforce(t, y, ivar, np.array([1.0]), **common)
    assert ref["power"][0] == pytest.approx(bf["power"][0], rel=1e-14)
    # Point 3 (phase 0.95) is lost to the copy, so the -1 never enters a box.
    y2 = y.copy()
    y2[3], y2[0] = 0.0, -1.0
    ref2 = oracle.bls_search(t, y2, ivar, np.array([1.0]), **common)
    assert ref2["depth"][0] > ref["depth"][0]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("obj_flag", [0, 1])
def test_oracle_matches_exact_bruteforce(seed, obj_flag):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(8, 40))
    t = np.sort(rng.uniform(0, 15, n)) + 2450000.0
    y = 1 + 0.01 * rng.standard_normal(n)
    y[rng.random(n) < 0.15] -= 0.04
    dy = rng.uniform(0.005, 0.02, n)
    periods = np.sort(rng.uniform(0.8, 4.0, 12))
    durations = np.array([0.1, 0.17, 0.3])
    per, bd, db, width = oracle.search_spec(periods, durations, 3)
    th, yh, ivh = oracle.host_inputs(t, y, dy)
    kw = dict(bin_duration=bd, dur_bins=db, oversample=3, width=width, obj_flag=obj_flag)
    ref = oracle.bls_search(th, yh, ivh, per, **kw)
    bf = bls_bruteforce(th, yh, ivh, per, **kw)
    finite = np.isfinite(bf["power"])
    assert np.array_equal(np.isfinite(ref["power"]), finite)
    np.testing.assert_allclose(ref["power"][finite], bf["power"][finite], rtol=1e-11)
    same = (ref["duration"] == bf["duration"]) & np.isclose(
        ref["transit_time"], bf["transit_time"], rtol=0, atol=1e-9
    )
    # Where the boxes differ, the oracle's box must be an exact tie of the best box
    # (boxes holding the same points, e.g. a bin and its wrap copy): FP64 prefix sums
    # order such ties by rounding, exact sums by first-wins.
    for p in np.flatnonzero(~same):
        assert _is_tie(th, yh, ivh, per[p], ref["duration"][p], ref["transit_time"][p],
                       bf["power"][p], obj_flag, **{k: kw[k] for k in ("bin_duration", "dur_bins", "oversample", "width")})
    assert same.mean() >= 0.75


def _is_tie(t, y, ivar, period, dur, tt, best, obj_flag, **kw):
    """Whether a box with this duration and transit time reaches ``best`` exactly."""
    rows, _ = box_table(t, y, ivar, period, **kw)
    s_w = math.fsum(np.asarray(ivar).tolist())
    s_y = math.fsum((np.asarray(y) * np.asarray(ivar)).tolist())
    t_min = float(np.min(t))
    bd = kw["bin_duration"]
    for di, n, yi, wi in rows:
        d = kw["dur_bins"][di]
        if d * bd != dur:
            continue
        if abs(math.fmod(n * bd + 0.5 * dur, period) + t_min - tt) > 1e-9:
            continue
        wo = s_w - wi
        depth = (s_y - yi) / wo - yi / wi
        obj = depth / math.sqrt(1 / wi + 1 / wo) if obj_flag == 0 else 0.5 * wi * depth**2
        if obj == pytest.approx(best, rel=1e-12):
            return True
    return False


def test_objectives_closed_form():
    # Weighted BLS (Kovacs et al. 2002) at the chosen box, from exact box sums.
    rng = np.random.default_rng(5)
    n = 30
    t = np.sort(rng.uniform(0, 10, n))
    y = rng.standard_normal(n) * 0.01
    y[(t % 2.0) < 0.2] -= 0.05
    ivar = rng.uniform(1e3, 2e4, n)
    per = np.array([2.0])
    kw = dict(bin_duration=0.05, dur_bins=[2, 4], oversample=2, width=44)
    ref = oracle.bls_search(t, y, ivar, per, obj_flag=0, **kw)
    rows, _ = box_table(t, y, ivar, 2.0, **{k: kw[k] for k in ("bin_duration", "dur_bins", "oversample", "width")})
    s_w = math.fsum(ivar.tolist())
    s_y = math.fsum((y * ivar).tolist())
    best = -math.inf
    for _di, _n, yi, wi in rows:
        wo = s_w - wi
        if wi <= 0 or wo <= 0:
            continue
        depth = (s_y - yi) / wo - yi / wi
        if depth < 0:
            continue
        best = max(best, depth / math.sqrt(1 / wi + 1 / wo))
    assert ref["power"][0] == pytest.approx(best, rel=1e-12)
    assert ref["depth_snr"][0] == ref["power"][0]
    assert ref["depth_err"][0] == pytest.approx(ref["depth"][0] / ref["power"][0], rel=1e-12)


@pytest.mark.parametrize("obj_flag", [0, 1])
def test_screen_is_exact_on_adversarial_cases(obj_flag):
    # tests/test_bls.py:397-424 of the reference: screened == unscreened bit for bit.
    periods = np.linspace(0.6, 9.0, 160)
    durations = np.array([0.02, 0.05, 0.11, 0.3])
    per, bd, db, width = oracle.search_spec(periods, durations, 3)
    for name, t, y, dy in adversarial_cases():
        ivar = 1.0 / dy**2
        kw = dict(bin_duration=bd, dur_bins=db, oversample=3, width=width, obj_flag=obj_flag)
        with np.errstate(all="ignore"):
            exact = oracle.bls_search(t, y, ivar, per, screen=False, **kw)
            fast = oracle.bls_search(t, y, ivar, per, **kw)
        for key in oracle.OUTPUT_NAMES:
            assert np.array_equal(fast[key], exact[key], equal_nan=True), (name, key)


def test_all_weights_underflow_gives_no_box():
    t, flux, _ = synthetic_eclipser(n=200)
    periods = np.linspace(0.6, 9.0, 50)
    with np.errstate(over="ignore"):
        out = oracle.bls_power(t, flux, np.full(t.size, 1e200), periods,
                               np.array([0.02, 0.05, 0.11, 0.3]), 3)
    assert np.all(np.isneginf(out["power"]))


def test_offset_and_scale_invariance():
    # tests/test_bls.py:182-210: power is invariant to a flux offset and to scaling
    # flux and errors together; a box swallowing every point is rejected.
    for seed in (1, 5):
        t, flux, err = two_runs(seed)
        periods = np.linspace(20.0, 180.0, 300)
        durations = np.array([0.2, 0.5, 1.0])
        base = oracle.bls_power(t, flux, err, periods, durations, 10, objective="likelihood")
        assert float(np.max(base["power"])) < 1e3
        for f2, e2 in ((flux + 1000.0, err), (flux * 1e12, err * 1e12)):
            other = oracle.bls_power(t, f2, e2, periods, durations, 10, objective="likelihood")
            np.testing.assert_allclose(other["power"], base["power"], rtol=1e-6, atol=1e-9)


def test_recovers_eclipser_period_and_depth():
    # tests/test_bls.py:73-79 recipe: P = 2.5 within 3e-3, depth 0.05 +- 0.01.
    t, flux, err = synthetic_eclipser(period=2.5, depth=0.05)
    periods = np.linspace(1.0, 5.0, 8000)
    out = oracle.bls_power(t, flux, err, periods, np.array([0.05, 0.1, 0.2]), 10)
    i = int(np.argmax(out["power"]))
    assert periods[i] == pytest.approx(2.5, rel=3e-3)
    assert out["depth"][i] == pytest.approx(0.05, abs=0.01)
    p = out["power"]
    assert (p[i] - p.mean()) / p.std() > 7.0


def test_threads_do_not_change_results():
    t, flux, err = synthetic_eclipser(n=500)
    periods = np.linspace(1.0, 5.0, 700)
    a = oracle.bls_power(t, flux, err, periods, np.array([0.05, 0.1]), 5, threads=1)
    b = oracle.bls_power(t, flux, err, periods, np.array([0.05, 0.1]), 5, threads=4)
    for k in oracle.OUTPUT_NAMES:
        assert np.array_equal(a[k], b[k])


def test_box_members_reproduce_the_oracle_power():
    # The tie proofs list a box's points by folding with the reference's formula; the
    # members of the oracle's own winning box must give back its power by direct sums.
    from oracle.ties import BoxInputs, box_members, recover_boxes

    rng = np.random.default_rng(1)
    t = np.sort(rng.uniform(0, 30, 300))
    y = 1 + 1e-3 * rng.standard_normal(300)
    y[np.mod(t, 2.3) < 0.1] -= 3e-3
    dy = np.full(300, 1e-3)
    periods = np.linspace(1.5, 3.0, 60)
    durations = np.array([0.05, 0.1])
    ref = oracle.bls_power(t, y, dy, periods, durations, 3)
    inp = BoxInputs.for_specs(t, [(periods, durations)], 3)
    _, yh, ivh = oracle.host_inputs(t, y, dy)
    yw = yh * ivh
    sy, si = math.fsum(yw), math.fsum(ivh)
    for i in range(periods.size):
        boxes = recover_boxes(ref, inp, i)
        assert boxes, i
        m = box_members(inp, i, *boxes[0])
        yin, ivin = math.fsum(yw[m]), math.fsum(ivh[m])
        depth = (sy - yin) / (si - ivin) - yin / ivin
        snr = depth / math.sqrt(1 / ivin + 1 / (si - ivin))
        assert snr == pytest.approx(ref["power"][i], rel=1e-9)


def test_tie_proof_refuses_a_box_with_other_points():
    # A planted mistake on the oracle's side: its box moved by one bin onto a different
    # point set must fail the proof; the unmoved box passes.
    from oracle.ties import BoxInputs, box_members, prove_same_points, recover_boxes

    rng = np.random.default_rng(2)
    t = np.sort(rng.uniform(0, 20, 400))
    y = 1 + 1e-3 * rng.standard_normal(400)
    dy = np.full(400, 1e-3)
    periods = np.array([1.7, 2.9])
    durations = np.array([0.1])
    ref = oracle.bls_power(t, y, dy, periods, durations, 4)
    inp = BoxInputs.for_specs(t, [(periods, durations)], 4)
    assert prove_same_points(ref, ref, inp, np.arange(2)) == 2
    n, d = recover_boxes(ref, inp, 0)[0]
    bd, P = float(inp.bin_duration[0]), float(inp.period[0])
    for shift in (1, 2, 3, -1, -2):
        n2 = n + shift
        if 0 <= n2 <= inp.n_bins(0) - d and not np.array_equal(
                box_members(inp, 0, n2, d), box_members(inp, 0, n, d)):
            break
    moved = {k: np.array(v, copy=True) for k, v in ref.items()}
    moved["transit_time"][0] = math.fmod(n2 * bd + 0.5 * d * bd, P) + inp.t_min
    with pytest.raises(AssertionError, match="different points"):
        prove_same_points(ref, moved, inp, np.array([0]))


def test_cross_period_tie_needs_the_same_points():
    # A curve whose deepest point is alone in the best box at many periods: those periods
    # tie in exact arithmetic (the same single point), and the oracle's FP64 powers there
    # agree to summation-order rounding; two periods whose boxes hold other points are
    # refused.
    from oracle.ties import BoxInputs, box_members, prove_cross_period_tie, recover_boxes

    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 40, 60))
    y = 1 + 1e-4 * rng.standard_normal(60)
    y[17] -= 0.05
    dy = np.full(60, 1e-4)
    periods = np.linspace(2.0, 3.0, 80)
    durations = np.array([0.01])
    ref = oracle.bls_power(t, y, dy, periods, durations, 3)
    inp = BoxInputs.for_specs(t, [(periods, durations)], 3)
    groups = {}
    for i in range(periods.size):
        boxes = recover_boxes(ref, inp, i)
        assert boxes, i
        groups.setdefault(tuple(box_members(inp, i, *boxes[0]).tolist()), []).append(i)
    single = groups.get((17,), [])
    assert len(single) >= 2
    assert prove_cross_period_tie(ref, inp, single[0], single[-1]) <= 64
    other = next(g[0] for k, g in groups.items() if k != (17,))
    with pytest.raises(AssertionError, match="ulp apart|different points"):
        prove_cross_period_tie(ref, inp, single[0], other)
