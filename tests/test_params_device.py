"""Parameters assembled on the device from a template row and the free
parameters of each set (ngpu_set_param_template / ngpu_set_genomes), against
the parameter matrix assembled on the host (the GA evaluators' path,
Figures/Figure4_DEAP/extractModel_mappings.py:158-171, :258-270)."""
import sys

import numpy as np
import pytest

from neurogpu_amd import NeuroGPUFromPkl, models
from neurogpu_amd._lib import NgpuError, Simulator, simulate_workload


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["active500", "ttpc1200", "passive64"])
def test_template_fills_to_the_host_matrix(name):
    """CPU: the sentinel template filled with the free parameters is the
    model's parameter matrix, entry for entry."""
    w = models.workload(name, nsets=7, nt=2)
    m = w.model
    tpl = m.param_template()
    nfree = len(m.free)
    for i in range(nfree):
        assert np.any(tpl == models.SENTINEL + i)
    np.testing.assert_array_equal(NeuroGPUFromPkl.allparams_from_mapping(tpl, w.x), m.allparams(w.x))


def _sim(w):
    m = w.model
    sim = Simulator(0)
    sim.set_tree(m.tree, m.mechset_id)
    sim.set_sim(m.v_init, m.rec_sites)
    sim.set_param_template(m.param_template(), m.ncomps, m.nparams, len(m.free))
    sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["active500", "ttpc1200"])
def test_device_assembly_equals_host_assembly(name):
    w = models.workload(name, nsets=6, nt=800)
    _say(f"case device-assembled parameters: {name} 6 sets, then 2")
    host, _ = simulate_workload(w)
    shape = host.shape
    with _sim(w) as sim:
        sim.set_genomes(w.x)
        out, st = sim.run()
        np.testing.assert_array_equal(out.reshape(shape), host)
        # the next generation: fewer sets, other values
        sim.set_genomes(w.x[[4, 1]])
        out2, _ = sim.run()
        np.testing.assert_array_equal(out2.reshape((2,) + shape[1:]), host[[4, 1]])
        f = sim.features(20.0, 180.0)
        assert f.shape == (2, shape[1], 5)


@pytest.mark.gpu
def test_template_errors():
    w = models.workload("active500", nsets=2, nt=10)
    m = w.model
    with Simulator(0) as sim:
        sim.set_tree(m.tree, m.mechset_id)
        with pytest.raises(NgpuError):
            sim.set_genomes(w.x)                             # no template yet
        tpl = m.param_template()
        with pytest.raises(NgpuError):
            sim.set_param_template(tpl, m.ncomps, m.nparams, len(m.free) + 1)   # free parameter without an entry
        with pytest.raises(NgpuError):
            sim.set_param_template(tpl, m.ncomps + 1, m.nparams, len(m.free))
        sim.set_param_template(tpl, m.ncomps, m.nparams, len(m.free))
        with pytest.raises(NgpuError):
            sim.run()                                       # no genomes yet
