"""Uneven batches: several morphologies of different size and step count in
one persistent launch through the device work queue.  Every member's traces
must equal those of the same context run alone (bitwise) and the FP64 oracle
(<= 1e-6 mV, identical spike counts), for every launch shape the autotuner can
choose."""
import sys

import numpy as np
import pytest

from conftest import TOL_MV, spike_counts
from neurogpu_amd import autotune, models
from neurogpu_amd._lib import MECH_MAINEN_GEN, MECH_TTPC_GEN, Batch, NgpuError, workload_simulator
from oracle import cpu_ref


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixed(nsets=(9, 5, 3, 2), nts=(300, 170, 90, 420), mechset="mainen"):
    """The uneven batch's four morphologies (the active cell's channel set by
    default), cut to a few sets and steps."""
    ws = models.mixed_workloads(1.0, 1.0, mechset=mechset)
    out = []
    for w, ns, nt in zip(ws, nsets, nts):
        out.append(models.Workload(w.name, w.model, w.x[:ns], w.amps[:, :nt], w.dts[:nt]))
    return out


def _shape(w):
    return (w.x.shape[0], w.amps.shape[0], len(w.model.rec_sites), w.dts.size)


def _alone(w, mechset=None):
    sim = workload_simulator(w, mechset=mechset)
    try:
        out, _ = sim.run(fetch=True)
    finally:
        sim.close()
    return out.reshape(_shape(w))


def _oracle(w, sets):
    m = w.model
    return cpu_ref.simulate(m.tree, m.mechset_id, m.allparams(w.x[sets]), w.amps, w.dts, m.stim_seg, m.stim_area,
                            m.v_init, m.rec_sites)


def test_candidates_cover_every_size():
    """CPU: the tuner's candidate shapes hold the tree and respect the rules."""
    for n, need in ((200, 1), (700, 2), (1500, 4), (3000, 8), (3072, 8)):
        cs = autotune.candidates(n)
        assert cs and min(c.waves_per_run for c in cs) == need
        for c in cs:
            assert 64 * c.waves_per_run * autotune.BATCH_NSLOT >= n
            assert c.waves_per_run == 1 or c.block_sync == 1
    assert autotune.candidates(3073) == []


def test_probe_repeats_sets_to_fill_the_device():
    """A member with fewer runs than the probe asks for is probed with its
    parameter sets repeated, in order, and its stimuli cut to the probe's
    steps."""
    w = models.workload("active500", nsets=5, nt=300)
    p = autotune._probe_workload(w, 40, 100)
    nst = w.amps.shape[0]
    assert p.x.shape[0] == -(-40 // nst) and p.nruns >= 40
    np.testing.assert_array_equal(p.x, w.x[np.arange(p.x.shape[0]) % 5])
    assert p.dts.size == 100 and p.amps.shape == (nst, 100)


def test_probe_size_covers_the_largest_item():
    """Items of one-wave groups hold up to 8 x 4 runs: a default probe of a
    small cell gives every CU four of them."""
    c1 = autotune.Choice(1, 0, 1)
    assert autotune.item_runs(c1) == 32 and autotune.item_runs(autotune.Choice(8, 1, 1)) == 6
    assert max(autotune.item_runs(c) for c in autotune.candidates(200)) == 32


def test_tree_key_is_stable_and_distinct():
    a, b = models.mixed_workloads(0.001, 0.001)[:2]
    assert autotune.tree_key(a) == autotune.tree_key(a)
    assert autotune.tree_key(a) != autotune.tree_key(b)


@pytest.mark.gpu
@pytest.mark.parametrize("mechset,choice", [("mainen", None), ("mainen", "wave_global"), ("mainen", "block_lds"),
                                            ("mainen", "two_waves"), ("ttpc", None), ("ttpc", "two_waves"),
                                            ("ttpc_gen", None), ("mainen_gen", None)])
def test_batch_matches_single_runs_and_oracle(mechset, choice):
    """ttpc_gen, mainen_gen: the cells through the mechanism sets generated
    from their NMODL files (MECH_TTPC_GEN, MECH_MAINEN_GEN), checked against
    the same oracle."""
    gen = {"ttpc_gen": MECH_TTPC_GEN, "mainen_gen": MECH_MAINEN_GEN}.get(mechset)
    ws = _mixed(mechset=mechset[:-len("_gen")] if gen else mechset)
    _say(f"case batch of {[w.model.tree.N for w in ws]} {mechset} segments, choice {choice}")
    alone = [_alone(w, gen) for w in ws]
    sims = [workload_simulator(w, mechset=gen) for w in ws]
    try:
        with Batch(0) as b:
            for i, s in enumerate(sims):
                b.add(s)
                if choice == "wave_global" and ws[i].model.tree.N <= 384:
                    b.set_launch(i, waves_per_run=1, block_sync=0, lds_staging=0)
                elif choice == "block_lds" and ws[i].model.tree.N <= 384:
                    b.set_launch(i, waves_per_run=1, block_sync=1, lds_staging=1)
                elif choice == "two_waves" and ws[i].model.tree.N <= 768:
                    b.set_launch(i, waves_per_run=2)
            st = b.run()
            infos = [b.member_info(i) for i in range(len(ws))]
            _say(f"  stats {st}")
        assert st["comp_steps"] == sum(w.comp_steps for w in ws)
        for w, s, a, info in zip(ws, sims, alone, infos):
            got = s.fetch().reshape(_shape(w))
            assert np.array_equal(got, a), f"{w.name}: batch differs from the single run"
            assert info["comp_steps"] == w.comp_steps
            assert 0.0 <= info["first_start_ms"] <= info["last_end_ms"]
            ref = _oracle(w, slice(0, 2))
            d = float(np.max(np.abs(got[:2] - ref)))
            _say(f"  {w.name}: max|batch-oracle| = {d:.3e} mV, shape {info}")
            assert d <= TOL_MV
            np.testing.assert_array_equal(spike_counts(got[:2]), spike_counts(ref))
    finally:
        for s in sims:
            s.close()


@pytest.mark.gpu
def test_batch_runs_twice_and_features_follow():
    ws = _mixed(nsets=(3, 2, 1, 1), nts=(200, 120, 60, 80))
    sims = [workload_simulator(w) for w in ws]
    try:
        with Batch(0) as b:
            for s in sims:
                b.add(s)
            b.run()
            first = [s.fetch().copy() for s in sims]
            b.run()
            for s, f in zip(sims, first):
                assert np.array_equal(s.fetch(), f)
        feats = sims[0].features([2.0], [4.0])
        assert feats.shape == (3, 1, 5)
    finally:
        for s in sims:
            s.close()


@pytest.mark.gpu
def test_batch_errors():
    w_t = _mixed(nsets=(2, 1, 1, 1), nts=(20, 20, 20, 20), mechset="ttpc")[0]
    w_m = models.workload("active500", nsets=2, nt=20)
    st = workload_simulator(w_t)
    sm = workload_simulator(w_m)
    try:
        with Batch(0) as b:
            assert b.add(st) == 0
            with pytest.raises(NgpuError):
                b.add(sm)                          # another mechanism set
            with pytest.raises(NgpuError):
                b.add(st)                          # twice
            with pytest.raises(NgpuError):
                b.set_launch(0, waves_per_run=3)
            with pytest.raises(NgpuError):
                b.set_launch(0, waves_per_run=2, block_sync=0)
            with pytest.raises(NgpuError):
                b.set_launch(5)
        with Batch(0) as b:
            with pytest.raises(NgpuError):
                b.run()                            # empty
    finally:
        st.close()
        sm.close()


@pytest.mark.gpu
def test_autotune_picks_a_working_shape(tmp_path):
    w = _mixed(nsets=(64, 1, 1, 1), nts=(60, 1, 1, 1))[0]
    cache = str(tmp_path / "tune.json")
    choice, rates = autotune.tune(w, probe_runs=512, probe_steps=40, cache=cache, log=_say)
    assert choice in autotune.candidates(w.model.tree.N)
    assert len(rates) == len(autotune.candidates(w.model.tree.N))
    again, _ = autotune.tune(w, cache=cache)
    assert again == choice
