"""Spike features: the CPU restatement pinned by known-answer traces, and the
on-device feature kernel against it."""
import math

import numpy as np
import pytest

from conftest import TOL_MV
from oracle import features_ref


def _triangles(centres, dt=0.1, nt=400, peak=30.0, slope=100.0, base=-70.0):
    t = np.arange(nt) * dt
    v = np.full(nt, base)
    for c in centres:
        m = np.abs(t - c) < (peak - base) / slope
        v[m] = np.maximum(v[m], peak - slope * np.abs(t[m] - c))
    return v, t


def test_known_answer_three_spikes():
    v, t = _triangles([10.0, 20.0, 32.0])
    f = features_ref.trace_features(v, t, t_on=5.0)
    # crossings of -30 at c -/+ 0.6: latency 10 - 0.6 - 5, width 1.2, ISI (10 + 12) / 2
    np.testing.assert_allclose(f, [3.0, 4.4, 100.0, 1.2, 11.0], rtol=0, atol=1e-9)


def test_no_spike_and_single_spike():
    v, t = _triangles([])
    f = features_ref.trace_features(v, t, 5.0)
    assert f[0] == 0 and all(math.isnan(x) for x in f[1:])
    v, t = _triangles([15.0])
    f = features_ref.trace_features(v, t, 5.0)
    assert f[0] == 1 and abs(f[1] - 9.4) < 1e-9 and math.isnan(f[4])


def test_spike_cut_by_the_end_counts_amplitude_not_width():
    v, t = _triangles([39.95], nt=400)
    f = features_ref.trace_features(v, t, 5.0)
    assert f[0] == 1 and math.isnan(f[3]) and f[2] > 0


def test_spikes_outside_the_window_are_ignored():
    # crossings of -30 at c - 0.6: 2.4 (before t_on = 5), 9.4 and 19.4 (in the
    # window), 35.4 (after t_off = 30)
    v, t = _triangles([3.0, 10.0, 20.0, 36.0])
    f = features_ref.trace_features(v, t, 5.0, 30.0)
    vbase = float(np.mean(v[t < 5.0]))          # the early spike raises the baseline
    np.testing.assert_allclose(f, [2.0, 4.4, 30.0 - vbase, 1.2, 10.0], rtol=0, atol=1e-9)
    # the whole trace as the window counts all four
    f = features_ref.trace_features(v, t, 0.0, 40.0)
    assert f[0] == 4 and abs(f[1] - 2.4) < 1e-9


def test_spike_that_starts_in_the_window_is_measured_in_full():
    # up at 19.4 < t_off = 20, down at 20.6: counted, width 1.2
    v, t = _triangles([20.0])
    f = features_ref.trace_features(v, t, 5.0, 20.0)
    np.testing.assert_allclose(f[:4], [1.0, 14.4, 100.0, 1.2], rtol=0, atol=1e-9)
    # up at 19.4 = t_off: outside
    f = features_ref.trace_features(v, t, 5.0, 19.4 - 1e-12)
    assert f[0] == 0 and math.isnan(f[2])


def test_baseline_is_mean_before_onset():
    v, t = _triangles([20.0])
    v[:50] = -60.0                      # t < 5 ms
    f = features_ref.trace_features(v, t, 5.0)
    assert abs(f[2] - (30.0 - (-60.0))) < 1e-9


def test_diverged_trace_has_every_feature_undefined():
    """A NaN sample anywhere in the trace (a diverged run) makes every
    feature NaN, the spike count included, whatever else the trace holds."""
    v, t = _triangles([10.0, 20.0])
    for i in (0, 150, 399):
        w = v.copy()
        w[i] = math.nan
        assert np.all(np.isnan(features_ref.trace_features(w, t, 5.0)))
    assert not np.any(np.isnan(features_ref.trace_features(v, t, 5.0)[:2]))


def test_diverged_runs_score_apart():
    """The evaluator's scores: an undefined feature difference costs PENALTY
    (the cap of a finite one), a diverged run DIVERGED_PENALTY on every
    objective of its stimulus, ten times more (PYS/efel_ext.py:73-83, :90-95
    score them 10000 and 100000)."""
    from neurogpu_amd import ga
    raw = np.array([[[3.0, 4.0, 90.0, 1.0, 11.0], [0.0] + [math.nan] * 4],
                    [[math.nan] * 5, [2.0, 5.0, 80.0, 1.1, math.nan]]])
    bad = ga.diverged(raw)
    assert bad.tolist() == [[False, False], [True, False]]
    filled = ga.fill_features(raw, 100.0)
    assert np.all(np.isfinite(filled)) and filled[1, 0, 0] == 0.0
    target = ga.Target(np.array([[3.0, 4.0, 90.0, 1.0, 11.0], [1.0, 50.0, 90.0, 1.0, 100.0]]),
                       np.array([1.0, 2.0, 5.0, 0.2, 2.0]))
    obj = ga.objectives(filled, target, bad).reshape(2, 2, len(ga.GA_FEATURES))
    assert np.all(obj[0, 0] == 0.0)
    assert np.all(obj[1, 0] == ga.DIVERGED_PENALTY) and ga.DIVERGED_PENALTY == 10 * ga.PENALTY
    assert np.all(obj[1, 1] < ga.PENALTY) and np.all(obj <= ga.DIVERGED_PENALTY)
    huge = filled.copy()
    huge[0, 1, 2] = 1e9
    assert ga.objectives(huge, target, bad).reshape(2, 2, -1)[0, 1, 2] == ga.PENALTY


@pytest.mark.gpu
def test_device_features_of_a_diverged_run():
    """A NaN parameter poisons its own run only: that run's device features
    are all NaN, the other runs' equal the restatement on their traces."""
    from neurogpu_amd import models
    from neurogpu_amd._lib import Simulator
    w = models.workload("active500", nsets=3, nt=1500)
    m = w.model
    P = w.params.copy()
    P[1, :] = np.nan
    with Simulator(0) as sim:
        sim.set_tree(m.tree, m.mechset_id)
        sim.set_sim(m.v_init, m.rec_sites)
        sim.set_params(P, m.ncomps, m.nparams)
        sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
        sim.run(fetch=False)
        gf = sim.features(20.0, 180.0)
        gv = sim.fetch().reshape(3, 3, 1, -1)
    assert np.all(np.isnan(gf[1])) and not np.any(np.isnan(gf[[0, 2], :, 0]))
    np.testing.assert_array_equal(gf, features_ref.features(gv, w.dts, 20.0, 180.0))


@pytest.mark.gpu
def test_device_features_match_oracle():
    from neurogpu_amd import models
    from neurogpu_amd._lib import Simulator
    from oracle import cpu_ref
    w = models.workload("active500", nsets=6, nt=2500)
    m = w.model
    with Simulator(0) as sim:
        sim.set_tree(m.tree, m.mechset_id)
        sim.set_sim(m.v_init, m.rec_sites)
        sim.set_params(w.params, m.ncomps, m.nparams)
        sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
        sim.run(fetch=False)
        gf = sim.features(20.0, 180.0)
        gv = sim.fetch().reshape(6, 3, 1, -1)
    ref = cpu_ref.simulate(m.tree, m.mechset_id, w.params, w.amps, w.dts, m.stim_seg, m.stim_area, m.v_init,
                           m.rec_sites)
    assert float(np.max(np.abs(gv - ref))) <= TOL_MV
    # the device features are a function of the trace: on the GPU's own traces
    # they equal the CPU restatement bit for bit; on the oracle's traces they
    # agree to the trace tolerance (spike counts exactly)
    np.testing.assert_array_equal(gf, features_ref.features(gv, w.dts, 20.0, 180.0))
    fr = features_ref.features(ref, w.dts, 20.0, 180.0)
    np.testing.assert_array_equal(gf[..., 0], fr[..., 0])
    np.testing.assert_allclose(gf, fr, rtol=0, atol=1e-5, equal_nan=True)
    assert np.nansum(gf[..., 0]) > 0
