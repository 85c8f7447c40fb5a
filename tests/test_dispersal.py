"""Branch points with more than two children: dispersed into chains of
extension compartments as the reference's model export does
(NeuroGPU_Base/runModel.hoc:158-215, hinesDisperseBranching), then scheduled
by the same map builder as every other tree (create_auxilliary_data_3.py:72-84)."""
import sys

import numpy as np
import pytest

from conftest import TOL_MV, spike_counts
from neurogpu_amd import formats, models, tree
from oracle.tree_ref import son_numbers
from oracle import cpu_ref


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_three_children_are_dispersed():
    #      0 -- 1 -- {2, 3, 4}      node 1 has three children
    par = np.array([-1, 0, 1, 1, 1])
    newpar, src = tree.disperse_branching(par)
    son = son_numbers(newpar)
    assert son.max() <= 2
    ext = src < 0
    assert ext.sum() == 2                                   # nchild - 1 extension nodes
    assert set(-src[ext] - 1) == {1}                         # both hang below node 1
    pos = {int(s): i for i, s in enumerate(src) if s >= 0}
    # node 1 keeps one child, the first extension; the children hang below extensions
    assert [i for i in range(len(newpar)) if newpar[i] == pos[1]] == [int(np.nonzero(ext)[0][0])]
    for c in (2, 3, 4):
        assert ext[newpar[pos[c]]]
    # children keep their order along the chain: child i on extension min(i, nchild - 2)
    xs = np.nonzero(ext)[0]
    assert newpar[pos[2]] == xs[0] and newpar[pos[3]] == xs[1] and newpar[pos[4]] == xs[1]
    tree.build_tree(newpar, np.ones(7), np.ones(7), np.ones(7), np.zeros(7, int), np.ones((1, 7), int), 1)


def test_root_with_two_children_is_dispersed():
    par = np.array([-1, 0, 0, 1, 2])
    newpar, src = tree.disperse_branching(par)
    assert son_numbers(newpar).max() <= 2 and (src < 0).sum() == 1


def test_trees_without_wide_branch_points_are_unchanged():
    m = models.workload("ttpc1200", nsets=1, nt=2).model
    newpar, src = tree.disperse_branching(m.morph.parent)
    np.testing.assert_array_equal(src, np.arange(m.tree.N))
    np.testing.assert_array_equal(newpar, m.morph.parent)


def test_branchy_model_maps_and_file_round_trip(tmp_path):
    m = models.branchy_active_model()
    raw = models.branchy()
    assert np.bincount(raw.parent[1:]).max() == 6
    t = m.tree
    assert t.son_no.max() <= 2 and t.N == raw.N + int(m.morph.ext.sum())
    # extension compartments: no mechanisms, the host's region
    assert t.bool_model[:, m.morph.ext].sum() == 0
    p = tmp_path / "BasicConstSegP.csv"
    formats.write_tree(str(p), t)
    t2 = formats.read_tree(str(p))
    formats.write_tree(str(tmp_path / "again.csv"), t2)
    assert open(p).read() == open(tmp_path / "again.csv").read()
    re = tree.rebuild_from(t2)
    for k in ("ks", "son_no", "seg_start", "seg_end", "fathers", "fidxs", "comp_by_level32"):
        np.testing.assert_array_equal(np.asarray(getattr(re, k)), np.asarray(getattr(t2, k)))


def test_branchy_solve_has_tiny_backward_error():
    """The extension nodes are as badly scaled as the reference's own
    exported trees; every solve mode of the oracle keeps a componentwise
    backward error below 1e-13 on the dispersed tree."""
    t = models.branchy_active_model().tree
    A0 = np.zeros((t.N, t.N))
    for h in range(t.N - 1):
        p = t.ks[h + 1] - 1
        A0[h, p] = t.f[h]
        A0[p, h] = t.e[h]
    rng = np.random.default_rng(4)
    u = np.abs(A0).sum(axis=0) + np.abs(A0).sum(axis=1) + rng.uniform(0.01, 1.0, t.N)
    A = A0 + np.diag(u)
    b = rng.normal(size=t.N)
    for mode in (0, 1, 2):
        x = cpu_ref.solve(t, u, b, mode)
        r = np.abs(A @ x - b)
        assert np.max(r / (np.abs(A) @ np.abs(x) + np.abs(b))) < 1e-13, mode


@pytest.mark.gpu
def test_branchy_cell_parity():
    from neurogpu_amd._lib import SOLVE_REFERENCE, simulate_workload
    m = models.branchy_active_model()
    x = m.sample(4, 1004)
    a, d = models.step_stimulus([0.2, 0.5, 1.0], 3000, 0.02, 10.0, 50.0)
    w = models.Workload("branchy", m, x, a, d)
    _say(f"case dispersed branchy cell: N={m.tree.N}, {w.nruns} runs x 3000 steps")
    gpu, st = simulate_workload(w)
    assert st["solve"] == SOLVE_REFERENCE                   # extension nodes: the reference's order
    ref = cpu_ref.simulate(m.tree, m.mechset_id, w.params, w.amps, w.dts, m.stim_seg, m.stim_area, m.v_init,
                           m.rec_sites)
    dmax = float(np.max(np.abs(gpu - ref)))
    _say(f"branchy: max|gpu-oracle| = {dmax:.3e} mV, spikes {spike_counts(ref).ravel().tolist()}")
    assert dmax <= TOL_MV
    np.testing.assert_array_equal(spike_counts(gpu), spike_counts(ref))
    assert spike_counts(ref).sum() > 0
