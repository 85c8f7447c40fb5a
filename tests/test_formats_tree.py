"""File formats and the tree-map scheduler against the reference's own data
files: the library's C++ scheduler and writer (neurogpu_amd.tree), and the
numpy restatement that checks it (oracle/tree_ref.py)."""
import os
import struct

import numpy as np
import pytest

from conftest import TREE_FILES, TTPC_DIR
from neurogpu_amd import formats, tree
from oracle import tree_ref


@pytest.mark.parametrize("name", sorted(TREE_FILES))
def test_tree_file_round_trip_is_byte_exact(name, tmp_path):
    t = formats.read_tree(TREE_FILES[name])
    out = tmp_path / "BasicConstSegP.csv"
    formats.write_tree(str(out), t)
    assert out.read_bytes() == open(TREE_FILES[name], "rb").read()


@pytest.mark.parametrize("name", sorted(TREE_FILES))
def test_library_writer_is_byte_exact(name, tmp_path):
    """ngpu_tree_write of the maps ngpu_tree_build computes from the parent
    vector alone: the reference generator's file, byte for byte."""
    t = formats.read_tree(TREE_FILES[name])
    out = tmp_path / "BasicConstSegP.csv"
    tree.write_tree(str(out), tree.rebuild_from(t))
    assert out.read_bytes() == open(TREE_FILES[name], "rb").read()


@pytest.mark.parametrize("impl", [tree, tree_ref], ids=["library", "oracle"])
@pytest.mark.parametrize("name", sorted(TREE_FILES))
def test_scheduler_reproduces_reference_maps(name, impl):
    """create_auxilliary_data_3 equivalent: every map rebuilt from the parent
    vector alone equals the reference generator's output."""
    t = formats.read_tree(TREE_FILES[name])
    r = impl.rebuild_from(t)
    for k in ("ks", "son_no", "rel_starts", "rel_ends", "rel_vec", "seg_start", "seg_end", "fathers",
              "fidxs", "comp_by_level32", "comp_by_flevel32", "lrel_starts", "lrel_ends",
              "flrel_starts", "flrel_ends", "ksb_m1"):
        np.testing.assert_array_equal(getattr(r, k), getattr(t, k), err_msg=k)
    assert (r.depth, r.logn_depth) == (t.depth, t.logn_depth)


def test_parent_vector_round_trip():
    t = formats.read_tree(TREE_FILES["ttpc512"])
    par = t.parent_segments()
    np.testing.assert_array_equal(tree.parent_to_ks(par), t.ks)
    assert np.all(par[1:] < np.arange(1, t.N))


def test_ancestor_table_definition():
    """FIdxs row k is the 2^k-th ancestor (1-based rows, N+1 beyond the root)."""
    t = formats.read_tree(TREE_FILES["mainen384"])
    N = t.N
    father = np.concatenate([t.ks[1:], [N + 1, N + 1]])
    for k in range(t.logn_depth):
        for r in range(1, N + 1, 7):
            a = r
            for _ in range(2 ** k):
                a = father[a - 1]
            assert t.fidxs[k, r - 1] == a
    # one more round would saturate every row: the table is complete
    last = t.fidxs[-1]
    nxt = np.array([father[a - 1] if a <= N else N + 1 for a in last])
    deep = last.copy()
    for _ in range(2 ** (t.logn_depth - 1)):
        deep = np.array([father[a - 1] if a <= N else N + 1 for a in deep])
    assert np.all(deep == N + 1)
    assert nxt.shape == last.shape


def _random_parent(rng, N, wide):
    """A random Hines-ordered parent vector; with ``wide`` some nodes get up
    to five children."""
    par = [-1]
    kids = [0]
    for s in range(1, N):
        limit = 5 if wide else 2
        while True:
            p = int(rng.integers(max(0, s - 12), s)) if rng.random() < 0.8 else int(rng.integers(0, s))
            if kids[p] < (limit if p else (limit - 1 if wide else 1)):
                break
        kids[p] += 1
        kids.append(0)
        par.append(p)
    return np.array(par)


@pytest.mark.parametrize("seed", range(6))
def test_library_scheduler_equals_the_oracle_on_random_trees(seed):
    """ngpu_tree_disperse and ngpu_tree_build against the numpy restatement on
    random morphologies (with wide branch points from seed 3 on): every map
    equal."""
    rng = np.random.default_rng(seed)
    N = int(rng.integers(2, 900))
    par = _random_parent(rng, N, wide=seed >= 3)
    p1, s1 = tree.disperse_branching(par)
    p2, s2 = tree_ref.disperse_branching(par)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(s1, s2)
    n = p1.size
    args = (p1, rng.normal(size=n), rng.normal(size=n), rng.uniform(0.5, 2, n), rng.integers(0, 3, n),
            rng.integers(0, 2, (2, n)), 3)
    a, b = tree.build_tree(*args), tree_ref.build_tree(*args)
    for k in ("ks", "son_no", "rel_starts", "rel_ends", "rel_vec", "seg_start", "seg_end", "fathers", "fidxs",
              "comp_by_level32", "comp_by_flevel32", "lrel_starts", "lrel_ends", "flrel_starts", "flrel_ends",
              "ksb_m1", "bool_model", "seg_to_comp"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert (a.depth, a.logn_depth, a.N) == (b.depth, b.logn_depth, b.N)


@pytest.mark.parametrize("impl", [tree, tree_ref], ids=["library", "oracle"])
def test_three_children_rejected(impl):
    par = np.array([-1, 0, 1, 1, 1])
    with pytest.raises(impl.TreeError):
        impl.build_tree(par, np.zeros(5), np.zeros(5), np.ones(5), np.zeros(5, int), np.ones((1, 5)), 1)


@pytest.mark.parametrize("impl", [tree, tree_ref], ids=["library", "oracle"])
def test_root_with_two_children_rejected(impl):
    """The generator numbers the root's children from 2, so a second child of
    the root would get SonNoVec 3 (create_auxilliary_data_3.py:72-84)."""
    par = np.array([-1, 0, 0])
    with pytest.raises(impl.TreeError):
        impl.build_tree(par, np.zeros(3), np.zeros(3), np.ones(3), np.zeros(3, int), np.ones((1, 3)), 1)


@pytest.mark.parametrize("bad", [[-1, 2, 0], [0, 0], [-1, -1], [-1, 0, 5]])
def test_non_hines_order_rejected(bad):
    with pytest.raises(tree.TreeError):
        tree.parent_to_ks(np.array(bad))
    with pytest.raises(tree.TreeError, match="Hines order|root"):
        tree.build_tree(np.array(bad), np.zeros(len(bad)), np.zeros(len(bad)), np.ones(len(bad)),
                        np.zeros(len(bad), int), np.ones((1, len(bad))), 1)
    with pytest.raises(tree.TreeError):
        tree.disperse_branching(np.array(bad))


def test_allparams_round_trip(tmp_path):
    t = formats.read_tree(TREE_FILES["ttpc512"])
    P = formats.read_allparams(os.path.join(TTPC_DIR, "AllParams.csv"), t.ncomps, 19)
    assert P.shape == (24, 19 * 241)
    out = tmp_path / "AllParams.csv"
    formats.write_allparams(str(out), P)
    np.testing.assert_array_equal(formats.read_allparams(str(out), t.ncomps, 19), P)
    np.testing.assert_array_equal(formats.quantise_params(P), P)


def test_allparams_wrong_count_is_an_error(tmp_path):
    p = tmp_path / "AllParams.csv"
    p.write_text("    2,\n1.0,2.0,\n3.0,\n")
    with pytest.raises(formats.FormatError):
        formats.read_allparams(str(p), 1, 2)
    p.write_text("    3,\n1.0,2.0,\n3.0,4.0,\n")
    with pytest.raises(formats.FormatError):
        formats.read_allparams(str(p), 1, 2)


def test_stim_and_sim_files():
    sim = formats.read_sim(os.path.join(TTPC_DIR, "Sim.csv"))
    assert sim.v_init == -65.0 and list(sim.rec_sites) == [0]
    st = formats.read_stim(TTPC_DIR)
    assert (st.nstim, st.nt, st.loc, st.stim_seg) == (1, 3168, 2, 1)
    assert st.amps.shape == (1, 3168) and np.all(st.dts == 0.1)
    assert st.dts.dtype == np.float64 and st.dts.max() < 1      # check_stims_and_times (GUI/Creating_Inputs.ipynb)


def test_stim_validation(tmp_path):
    (tmp_path / "Stim_meta.csv").write_text("1\n4\n0\n1\n0.0\n")
    (tmp_path / "Stim_raw.csv").write_text("1,1,1,1,\n")
    (tmp_path / "times.csv").write_text("0.1,0.1,0.1,0.1,\n")
    with pytest.raises(formats.FormatError):
        formats.read_stim(str(tmp_path))
    (tmp_path / "Stim_meta.csv").write_text("1\n4\n0\n1\n10.0\n")
    (tmp_path / "times.csv").write_text("0.1,0.0,0.1,0.1,\n")
    with pytest.raises(formats.FormatError):
        formats.read_stim(str(tmp_path))


def test_vhot_file_layout(tmp_path):
    v = np.arange(12, dtype=np.float64).reshape(2, 1, 1, 6)
    p = tmp_path / "VHotP0.dat"
    formats.write_vhot(str(p), v)
    raw = p.read_bytes()
    assert struct.unpack("<ii", raw[:8]) == (12, 3)
    np.testing.assert_array_equal(formats.read_vhot(str(p)), v.ravel())
