"""GPU parity: libneurogpu's HIP kernels against the FP64 oracle on the same
inputs.  The bar is max-abs trace difference <= 1e-6 mV and identical spike
counts.  The solve keeps the reference's elimination arithmetic bit for bit;
the mechanism code rearranges the rate equations, so traces differ from the
oracle by rounding (about 1e-8 mV at most over 10,000 steps)."""
import os
import sys

import numpy as np
import pytest

from conftest import TOL_MV, TTPC_DIR, spike_counts
from neurogpu_amd import formats, models
from neurogpu_amd._lib import MECH_MAINEN_GEN, MECH_TTPC_GEN, SOLVE_FAST, SOLVE_REFERENCE, NgpuError, Simulator, simulate_workload
from oracle import cpu_ref

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle(w, sets=None):
    m = w.model
    x = w.x if sets is None else w.x[sets]
    return cpu_ref.simulate(m.tree, m.mechset_id, m.allparams(x), w.amps, w.dts, m.stim_seg, m.stim_area,
                            m.v_init, m.rec_sites)


def _check(gpu, ref, what):
    d = float(np.max(np.abs(gpu - ref)))
    _say(f"{what}: max|gpu-oracle| = {d:.3e} mV")
    assert np.all(np.isfinite(gpu))
    assert d <= TOL_MV, what
    np.testing.assert_array_equal(spike_counts(gpu), spike_counts(ref))
    return d


def test_passive64_full_size():
    w = models.workload("passive64")
    _say(f"case passive64 full: {w.nruns} runs x {w.dts.size} steps")
    gpu, st = simulate_workload(w)
    _check(gpu, _oracle(w), "passive64")


@pytest.mark.parametrize("launch", [None, dict(waves_per_run=2), dict(runs_per_group=2), dict(waves_per_run=16),
                                    dict(waves_per_run=1, runs_per_group=4), dict(waves_per_run=2, runs_per_group=3),
                                    dict(waves_per_run=4, runs_per_group=6),
                                    dict(steps_per_launch=700), dict(steps_per_launch=500, use_graph=True)])
def test_active500_small(launch):
    w = models.workload("active500", nsets=6, nt=1500)
    _say(f"case active500 small launch={launch}")
    gpu, st = simulate_workload(w, launch=launch)
    _check(gpu, _oracle(w), f"active500 {launch}")


@pytest.mark.parametrize("launch", [None, dict(waves_per_run=4), dict(waves_per_run=8, runs_per_group=3),
                                    dict(waves_per_run=16, runs_per_group=4)])
def test_ttpc1200_small(launch):
    w = models.workload("ttpc1200", nsets=3, nt=1200)
    _say(f"case ttpc1200 small launch={launch}")
    gpu, st = simulate_workload(w, launch=launch)
    assert st["solve"] == SOLVE_FAST                 # a generated tree: the reciprocal form by default
    _check(gpu, _oracle(w), f"ttpc1200 {launch}")


@pytest.mark.parametrize("launch", [None, dict(waves_per_run=4), dict(waves_per_run=16, runs_per_group=4)])
def test_ttpc1200_generated_mechanisms_small(launch):
    """The same cell through the mechanism set generated from its NMODL files
    (MECH_TTPC_GEN, gen_ttpc.h) instead of the hand-written one: the oracle's
    bar."""
    w = models.workload("ttpc1200", nsets=3, nt=1200)
    _say(f"case ttpc1200 generated mechanisms launch={launch}")
    gpu, st = simulate_workload(w, launch=launch, mechset=MECH_TTPC_GEN)
    _check(gpu, _oracle(w), f"ttpc1200 generated {launch}")


def test_ttpc1200_generated_mechanisms_full_protocol():
    """Configuration 3's whole protocol with the generated mechanism set: the
    oracle's bar, and within rounding of the hand-written set's traces."""
    w = models.workload("ttpc1200", nsets=2)
    _say(f"case ttpc1200 generated mechanisms, full protocol: {w.nruns} runs x {w.dts.size} steps")
    gen, _ = simulate_workload(w, mechset=MECH_TTPC_GEN)
    hand, _ = simulate_workload(w)
    _check(gen, _oracle(w), "ttpc1200 generated full protocol")
    d = float(np.max(np.abs(gen - hand)))
    _say(f"ttpc1200: max|generated - hand-written| = {d:.3e} mV")
    assert d <= TOL_MV


@pytest.mark.parametrize("launch", [None, dict(waves_per_run=1, runs_per_group=4), dict(waves_per_run=4, runs_per_group=3)])
def test_active500_generated_mechanisms_small(launch):
    """The Mainen cell through the mechanism set generated from its NMODL
    files (MECH_MAINEN_GEN, gen_mainen.h): the oracle's bar."""
    w = models.workload("active500", nsets=6, nt=1500)
    _say(f"case active500 generated mechanisms launch={launch}")
    gpu, st = simulate_workload(w, launch=launch, mechset=MECH_MAINEN_GEN)
    _check(gpu, _oracle(w), f"active500 generated {launch}")


def test_active500_generated_mechanisms_full_protocol():
    """Configuration 2's whole 10,000-step protocol with the generated Mainen
    set: the oracle's bar, and within rounding of the hand-written set."""
    w = models.workload("active500", nsets=4)
    _say(f"case active500 generated mechanisms, full protocol: {w.nruns} runs x {w.dts.size} steps")
    gen, _ = simulate_workload(w, mechset=MECH_MAINEN_GEN)
    hand, _ = simulate_workload(w)
    _check(gen, _oracle(w), "active500 generated full protocol")
    d = float(np.max(np.abs(gen - hand)))
    _say(f"active500: max|generated - hand-written| = {d:.3e} mV")
    assert d <= TOL_MV


@pytest.mark.parametrize("name", ["active500", "ttpc1200", "passive64"])
def test_both_eliminations_agree(name):
    """The reference's operation order and the reciprocal form on the
    same generated tree: both within the bar of the oracle, and within
    rounding of each other."""
    w = models.workload(name, nsets=3, nt=1500)
    _say(f"case {name}: reference-order vs reciprocal-form elimination")
    ref_order, st1 = simulate_workload(w, solve=SOLVE_REFERENCE)
    fast, st2 = simulate_workload(w, solve=SOLVE_FAST)
    assert st1["solve"] == SOLVE_REFERENCE and st2["solve"] == SOLVE_FAST
    oracle = _oracle(w)
    _check(ref_order, oracle, f"{name} reference order")
    _check(fast, oracle, f"{name} reciprocal form")
    d = float(np.max(np.abs(fast - ref_order)))
    _say(f"{name}: max|fast - reference order| = {d:.3e} mV")
    assert d <= 1e-8


def test_ttpc1200_full_protocol():
    """Configuration 3 over its whole 300 ms protocol (12,000 steps), two
    parameter sets x all five stimuli, every run checked."""
    w = models.workload("ttpc1200", nsets=2)
    _say(f"case ttpc1200 full protocol: {w.nruns} runs x {w.dts.size} steps")
    gpu, st = simulate_workload(w)
    _check(gpu, _oracle(w), "ttpc1200 full protocol")


@pytest.mark.parametrize("name,cuts", [("active500", (0, 500, 1100, 1500)), ("ttpc1200", (0, 1, 64, 700, 1200))])
def test_windows_equal_the_whole_run(name, cuts):
    """ngpu_run_steps over consecutive windows carries the state: the traces
    equal one ngpu_run of the same steps bit for bit, and a window that does
    not start where the last one ended is refused."""
    w = models.workload(name, nsets=4, nt=cuts[-1])
    _say(f"case {name} windows {cuts}")
    whole, _ = simulate_workload(w)
    m = w.model
    with Simulator(0) as sim:
        sim.set_tree(m.tree, m.mechset_id)
        sim.set_sim(m.v_init, m.rec_sites)
        sim.set_params(w.params, m.ncomps, m.nparams)
        sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
        with pytest.raises(NgpuError):
            sim.run_steps(cuts[1], cuts[2])        # no state at cuts[1] yet
        for a, b in zip(cuts[:-1], cuts[1:]):
            st = sim.run_steps(a, b)
            assert st["comp_steps"] == w.nruns * m.tree.N * (b - a)
        with pytest.raises(NgpuError):
            sim.run_steps(cuts[1], cuts[2])        # the state is at the end now
        got = sim.fetch().reshape(whole.shape)
        feats = sim.features(20.0, 180.0)
        assert feats.shape == (4, w.amps.shape[0], 5)
    np.testing.assert_array_equal(got, whole)


def test_open_reads_sizes_from_the_files():
    """A context from ngpu_open knows its sizes (NSets, NStimuli, Nt from the
    files), so features() and fetch() size their buffers from the C side."""
    _say("case ngpu_open sizes")
    with Simulator.open(TTPC_DIR) as sim:
        assert (sim.nsets, sim.nstim, sim.nrec, sim.nt) == (24, 1, 1, 3168)
        sim.run(fetch=False)
        f = sim.features(10.0, 70.0)
        assert f.shape == (24, 1, 5)
        assert sim.fetch().size == 24 * 3168


@pytest.mark.parametrize("mechset", [-1, MECH_TTPC_GEN])
def test_reference_ttpc_fixture_through_file_boundary(tmp_path, mechset):
    """The reference's own data directory, parsed by ngpu_open (C++), run, and
    written as VHotP.dat; compared with the oracle run on the Python parse.
    mechset -1 infers the hand-written TTPC set from the tree file; the
    generated set reads the same files."""
    _say(f"case ttpc512 fixture via ngpu_open, mechset {mechset}")
    sim = Simulator.open(TTPC_DIR, mechset=mechset)
    out = tmp_path / "VHotP0.dat"
    _, st = sim.run(fetch=False)
    assert st["solve"] == SOLVE_REFERENCE            # zero-area nodes: the reference's order
    sim.write_vhot(str(out))
    sim.close()
    t = formats.read_tree(os.path.join(TTPC_DIR, "BasicConstSegP.csv"))
    P = formats.read_allparams(os.path.join(TTPC_DIR, "AllParams.csv"), t.ncomps, 19)
    sd = formats.read_sim(os.path.join(TTPC_DIR, "Sim.csv"))
    st = formats.read_stim(TTPC_DIR)
    ref = cpu_ref.simulate(t, 1, P, st.amps, st.dts, st.stim_seg, st.area, sd.v_init, sd.rec_sites)
    gpu = formats.read_vhot(str(out)).reshape(ref.shape)
    _check(gpu, ref, "ttpc512 fixture")
    assert np.all(np.abs(gpu) < 1000)          # Code/test_output.py bound


def test_active_full_length_sampled():
    """Configuration 1 at full length (10,000 steps), 16 parameter sets, all
    three stimuli, every run checked."""
    w = models.workload("active500", nsets=16)
    _say("case active500 full length, 16 sets")
    gpu, st = simulate_workload(w)
    _check(gpu, _oracle(w), "active500 full length")


def test_determinism_and_identical_instances():
    w = models.workload("active500", nsets=8, nt=800)
    w.x[4:] = w.x[:4]
    _say("case determinism")
    a, _ = simulate_workload(w)
    b, _ = simulate_workload(w)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a[:4], a[4:])


def test_error_paths():
    w = models.workload("active500", nsets=2, nt=10)
    m = w.model
    with Simulator(0) as sim:
        with pytest.raises(NgpuError):
            sim.run()                                   # nothing set yet
        sim.set_tree(m.tree, m.mechset_id)
        with pytest.raises(NgpuError):
            sim.set_params(w.params, m.ncomps + 1, m.nparams)
        with pytest.raises(NgpuError):
            sim.set_params(w.params, m.ncomps, m.nparams - 1)
        bad = w.dts.copy()
        bad[3] = 0.0
        with pytest.raises(NgpuError):
            sim.set_stimuli(w.amps, bad, m.stim_seg, m.stim_area)
        with pytest.raises(NgpuError):
            sim.set_stimuli(w.amps, w.dts, m.tree.N, m.stim_area)
        with pytest.raises(NgpuError):
            sim.set_sim(m.v_init, [m.tree.N])
        with pytest.raises(NgpuError):
            sim.set_launch(waves_per_run=9)


def test_nan_parameter_poisons_only_its_run():
    w = models.workload("active500", nsets=3, nt=300)
    P = w.params
    P[1, :] = np.nan
    m = w.model
    with Simulator(0) as sim:
        sim.set_tree(m.tree, m.mechset_id)
        sim.set_sim(m.v_init, m.rec_sites)
        sim.set_params(P, m.ncomps, m.nparams)
        sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
        out, _ = sim.run()
    out = out.reshape(3, w.amps.shape[0], 1, w.dts.size)
    assert np.all(np.isfinite(out[0])) and np.all(np.isfinite(out[2]))
    assert not np.all(np.isfinite(out[1]))
