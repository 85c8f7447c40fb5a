"""ngpu_features_noalloc (the feature entry point that allocates nothing and
stages through a fixed device array) against ngpu_features, through the C
ABI: the same features bit for bit on two calls in a row, the second with a
smaller input, and sizes beyond the fixed array refused with an error."""
import ctypes as C
import sys

import numpy as np
import pytest

from neurogpu_amd import models
from neurogpu_amd._lib import Simulator, lib, workload_simulator

pytestmark = pytest.mark.gpu

NGPU_E_ARG = -1
_DP = C.POINTER(C.c_double)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _features(L, name, sim, on, off):
    out = np.full((sim.nsets, sim.nstim, 5), -1.0)
    rc = getattr(L, name)(sim._h, on.ctypes.data_as(_DP), off.ctypes.data_as(_DP), -30.0, out.ctypes.data_as(_DP))
    return rc, out


def test_noalloc_features_equal_the_allocating_entry_point():
    L = lib()
    # the second call is smaller; its 1,300 steps (26 ms) end soon after the
    # stimulus window opens, so only the first is required to spike
    for nsets, nt in ((12, 1500), (3, 1300)):
        w = models.workload("active500", nsets=nsets, nt=nt)
        _say(f"case no-alloc features: active500 {nsets} sets x {w.amps.shape[0]} stimuli x {nt} steps")
        sim = workload_simulator(w)
        try:
            sim.run(fetch=False)
            on = np.array([20.0, 22.5, 25.0][:sim.nstim])
            off = np.array([180.0, 150.0, 120.0][:sim.nstim])
            rc0, want = _features(L, "ngpu_features", sim, on, off)
            rc1, got = _features(L, "ngpu_features_noalloc", sim, on, off)
            assert rc0 == 0 and rc1 == 0, (rc0, rc1, L.ngpu_last_error(sim._h))
            np.testing.assert_array_equal(got, want)
            if nsets == 12:
                assert np.nansum(got[..., 0]) > 0
        finally:
            sim.close()


def test_noalloc_refuses_sizes_beyond_its_array():
    L = lib()
    w = models.workload("passive64", nsets=1000, nt=4)
    m = w.model
    amps = np.repeat(w.amps, 250, axis=0)              # 1,000 sets x 250 stimuli = 250,000 runs
    _say("case no-alloc features: 250,000 runs (more than the fixed array holds)")
    with Simulator(0) as sim:
        sim.set_tree(m.tree, m.mechset_id)
        sim.set_sim(m.v_init, m.rec_sites)
        sim.set_params(w.params, m.ncomps, m.nparams)
        sim.set_stimuli(amps, w.dts, m.stim_seg, m.stim_area)
        on, off = np.zeros(250), np.full(250, 0.1)
        rc, _ = _features(L, "ngpu_features_noalloc", sim, on, off)
        assert rc == NGPU_E_ARG
        assert b"scratch" in L.ngpu_last_error(sim._h)
