"""The mechanism-pass schedule (neurogpu_amd/csrc/schedule.h), compiled for
the host: for the workloads' trees and every group size, each (chunk of 64
positions, run of the group) is covered by exactly one segment of the queue,
segments of a chunk with run-time-mask positions cover one run, and the queue
is ordered by estimated cost, longest first (the waves take segments from it
in that order)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from neurogpu_amd import models

_IP = C.POINTER(C.c_int)
FIXED = {1: {0x408, 0x1618}, 2: {0x6F}, 0: {0, 1}}


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("sched") / "libschedule_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "neurogpu_amd", "csrc"),
                    os.path.join(ROOT, "tests", "schedule_host.cpp"), "-o", out], check=True)
    L = C.CDLL(out)
    L.host_mech_schedule.restype = C.c_int
    L.host_mech_schedule.argtypes = [_IP, C.c_int, C.c_int, C.c_int, C.c_int, _IP, C.c_int]
    L.host_choose_shape.restype = None
    L.host_choose_shape.argtypes = [C.c_long, C.c_int, _IP, C.c_int, C.c_int, C.c_double, C.c_double, _IP]
    return L


def _positions_masks(name):
    m = models.workload(name, nsets=1, nt=10).model
    t = m.tree
    bm = np.asarray(t.bool_model).reshape(-1, t.N)
    masks = (bm.astype(np.int64) << np.arange(bm.shape[0])[:, None]).sum(0)
    return np.sort(masks, kind="stable").astype(np.int32), t.N, m.mechset_id


def _schedule(lib, qmask, N, ms, R):
    NQ = (N + 63) // 64 * 64
    out = np.zeros(4096, np.int32)
    n = lib.host_mech_schedule(qmask.ctypes.data_as(_IP), N, NQ, ms, R, out.ctypes.data_as(_IP), out.size)
    assert n > 0
    return out[:n], NQ


def _cost(qmask, N, k):
    ms = set(int(x) for x in qmask[64 * k:min(64 * k + 64, N)])
    return sum(bin(m).count("1") + 1 for m in ms)


@pytest.mark.parametrize("name", ["passive64", "active500", "ttpc1200"])
@pytest.mark.parametrize("R", [1, 2, 3, 4, 6, 8])
def test_every_position_and_run_once(lib, name, R):
    """Chunk segments cover each (chunk, run) with specialised-mask positions
    once, in segments of up to three runs; the packed segments cover the
    (run-time-mask position, run) pairs, 64 a segment; the queue is ordered
    by estimated cost, longest first."""
    qmask, N, ms = _positions_masks(name)
    s, NQ = _schedule(lib, qmask, N, ms, R)
    assert s[0] == s.size - 1
    fixed = np.array([int(m) in FIXED[ms] for m in qmask])
    ngen = int((~fixed).sum())
    covered = np.zeros((NQ // 64, R), np.int32)
    packed = []
    costs = []
    gcost = 3 * sum(bin(int(m)).count("1") + 1 for m in set(qmask[~fixed].tolist()))
    for d in s[1:]:
        c, r0, n = d >> 8, (d >> 4) & 15, d & 15
        if n == 0:
            packed.append(c)
            costs.append(gcost)
            continue
        assert 0 <= c < NQ // 64 and r0 + n <= R and n <= 3
        assert fixed[64 * c:min(64 * c + 64, N)].any()
        covered[c, r0:r0 + n] += 1
        costs.append(_cost(qmask[fixed], N, c) if False else
                     sum(bin(int(m)).count("1") + 1 for m in set(qmask[64 * c:min(64 * c + 64, N)][fixed[64 * c:min(64 * c + 64, N)]].tolist())) * n)
    has_fixed = np.array([fixed[64 * k:min(64 * k + 64, N)].any() for k in range(NQ // 64)])
    np.testing.assert_array_equal(covered[has_fixed], 1)
    np.testing.assert_array_equal(covered[~has_fixed], 0)
    assert sorted(packed) == list(range((ngen * R + 63) // 64))
    assert all(a >= b for a, b in zip(costs, costs[1:]))      # longest first


def test_runtime_mask_positions_are_packed(lib):
    """ttpc1200's 21 soma and axon segments (run-time-mask code) over a group's
    six runs: two packed segments of 64 pairs, at the head of the queue, in
    place of six single-run passes over their chunk."""
    qmask, N, ms = _positions_masks("ttpc1200")
    s, NQ = _schedule(lib, qmask, N, 1, 6)
    assert int(sum(int(m) not in FIXED[1] for m in qmask)) == 21
    assert [d & 15 for d in s[1:3]] == [0, 0] and all(d & 15 for d in s[3:])


def _shape(lib, nruns, ws, run_kib, table_kib, r_only=0, ncu=256):
    w = np.asarray(ws, dtype=np.int32)
    out = np.zeros(2, dtype=np.int32)
    lib.host_choose_shape(nruns, ncu, w.ctypes.data_as(_IP), w.size, r_only, run_kib, table_kib,
                          out.ctypes.data_as(_IP))
    return int(out[0]), int(out[1])


# LDS per run and per workgroup (KiB) of the 500- and 1,200-segment cells
A500 = dict(run_kib=8.4, table_kib=13.0)
T1200 = dict(run_kib=19.4, table_kib=29.0)


@pytest.mark.parametrize("nruns,expect", [
    (3072, (2, 3)),     # 1,024 sets x 3 stimuli: 256 workgroups of 12 runs, one per CU
    (24576, (1, 2)),    # 8,192 sets x 3: whole rounds of 16-run workgroups
    (16, (4, 1)),       # a handful of runs: spread over CUs
])
def test_shape_small_cell(lib, nruns, expect):
    """Shapes chosen for the 500-segment cell.  On one MI355X, 3,072 runs
    (2,000 steps) took 253 ms with groups of 2 waves x 3 runs and 314 ms with
    1 wave x 2 runs (192 workgroups, 64 CUs idle)."""
    assert _shape(lib, nruns, [1, 2, 4], **A500) == expect


@pytest.mark.parametrize("nruns,expect", [(40960, 6), (81920, 6), (5120, 5), (5, 1)])
def test_shape_large_cell(lib, nruns, expect):
    """16-wave groups of the 1,200-segment cell: six runs at full size; 5,120
    runs (a population of 1,024 x 5 stimuli) fill four rounds of 1,024
    workgroups with five runs instead of 854 workgroups of six."""
    assert _shape(lib, nruns, [16], **T1200) == (16, expect)


def test_shape_fixed_runs_and_no_fit(lib):
    assert _shape(lib, 40960, [16], r_only=4, **T1200) == (16, 4)
    assert _shape(lib, 100, [16], run_kib=200.0, table_kib=0.0)[1] == 0
    # with fewer CUs the same runs fill more rounds: the largest workgroups win
    assert _shape(lib, 3072, [1, 2, 4], ncu=8, **A500) == (1, 2)


def test_shape_minimises_rounds_times_size(lib):
    """Against a brute-force evaluation on random sizes: no other fitting
    shape with full workgroups has fewer rounds x (6 + runs)."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        nruns = int(rng.integers(1, 100000))
        ncu = int(rng.integers(1, 300))
        W, R = _shape(lib, nruns, [1, 2, 4], ncu=ncu, **A500)
        def groups(w, r):
            if r * A500["run_kib"] + A500["table_kib"] > 160:
                return 0
            g = 1
            while (g + 1) * w <= 8 and (g + 1) * r * A500["run_kib"] + A500["table_kib"] <= 160:
                g += 1
            return g
        def cost(w, r):
            per = groups(w, r) * r
            return -(-(-(-nruns // per)) // ncu) * (6 + per)
        cands = [(w, r) for w in (1, 2, 4) for r in range(1, (4 if w == 1 else 6) + 1) if groups(w, r) * w >= 8]
        assert (W, R) in cands
        assert cost(W, R) == min(cost(w, r) for w, r in cands)
