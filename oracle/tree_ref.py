"""Tree-map scheduler restated in numpy: the checker of the library's C++
scheduler (neurogpu_amd/csrc/treemap.h, ngpu_tree_build, ngpu_tree_disperse).

Test infrastructure only: imported by tests/, never by the product path.

It computes the arrays of the reference's generator
(NeuroGPU_Base/scripts/pyscripts/create_auxilliary_data_3.py:22-315 and
make_tree_from_parent_vec.py:16-59) and the dispersal of wide branch points
(hinesDisperseBranching, NeuroGPU_Base/runModel.hoc:158-215).  Pinned by
integer equality with the maps of the reference's own tree files
(tests/golden/trees, tests/golden/ttpc512; tests/test_formats_tree.py) and by
the definitions its tests check (ancestor distances, son numbering).

Conventions (all as in the file): row h = N-1-s for segment s; ``Ks[r]`` is the
1-based parent row of 1-based row r (r = 1..N-1; ``Ks[0]`` = 0); run and father
tables are 1-based.
"""
from __future__ import annotations

import math
from typing import Dict, List

import numpy as np

from neurogpu_amd.formats import TreeData


class TreeError(ValueError):
    """The parent vector cannot be scheduled by the reference's solver."""


def parent_to_ks(par: np.ndarray) -> np.ndarray:
    """Ks from a segment parent vector (root -1); inverse of
    ``parent(s) = N - Ks[N-s]`` (``create_auxilliary_data_3.py:45-50``)."""
    par = np.asarray(par, dtype=np.int64)
    N = par.size
    if N < 1 or par[0] != -1:
        raise TreeError("segment 0 must be the root (parent -1)")
    if np.any(par[1:] < 0) or np.any(par[1:] >= np.arange(1, N)):
        raise TreeError("parent vector must satisfy 0 <= parent(s) < s (Hines order)")
    ks = np.zeros(N, dtype=np.int64)
    for j in range(1, N):
        ks[j] = N - par[N - j]
    return ks


def son_numbers(par: np.ndarray) -> np.ndarray:
    """SonNoVec (``create_auxilliary_data_3.py:72-84``): the ordinal of each
    segment among its parent's children in segment order.  The root is counted
    as a child of itself first, so the root's children are numbered from 2."""
    N = par.size
    used = np.zeros(N, dtype=np.int64)
    son = np.zeros(N, dtype=np.int64)
    for s in range(N):
        p = 0 if s == 0 else int(par[s])
        used[p] += 1
        son[s] = used[p]
    son[0] = 0
    return son


def ancestor_table(ks: np.ndarray):
    """FIdxs and LognDepth (``create_auxilliary_data_3.py:53-67``): row k holds
    the ancestor at distance 2^k of every row (1-based, N+1 = beyond the root);
    rows stop before the first one that is fully saturated."""
    N = ks.size
    father = np.concatenate([ks[1:], [N + 1, N + 1]])          # rows 1..N+1
    K = int(math.ceil(math.log2(N))) + 3 if N > 1 else 3
    rows = []
    cur = np.arange(1, N + 2)
    for k in range(K):
        step = cur
        for _ in range(2 ** k):
            step = father[step - 1]
        rows.append(step)
    rows = np.array(rows)
    last = rows[-1]
    k0 = next(k for k in range(K) if np.array_equal(rows[k], last))
    return rows[:k0, :N].astype(np.int64), int(k0)


def _cell_2_vec(groups: List[List[int]]):
    starts, ends, vec, pos = [], [], [], 0
    for g in groups:
        starts.append(pos)
        ends.append(pos + len(g) - 1)
        vec.extend(g)
        pos += len(g)
    return np.array(starts, np.int64), np.array(ends, np.int64), np.array(vec, np.int64)


def _pad32(items: List[int]) -> np.ndarray:
    n = int(math.ceil(len(items) / 32.0)) * 32
    out = np.full(n, items[-1], dtype=np.int64)
    out[: len(items)] = items
    return out.reshape(-1, 32)


def schedule(ks: np.ndarray) -> Dict[str, np.ndarray]:
    """Runs, fathers, levels and the 32-wide level tables
    (``make_tree_from_parent_vec.py:17-59``, ``create_auxilliary_data_3.py:165-206``)."""
    N = ks.size
    rows = np.arange(N)
    fathers = np.unique(ks[(ks != rows + 1) & (ks > 0)])
    related = [list(np.nonzero(ks == fa)[0]) for fa in fathers]
    rs, re, rv = _cell_2_vec(related) if fathers.size else (np.zeros(0, np.int64),) * 3
    is_f = np.zeros(N + 2, dtype=bool)
    is_f[fathers] = True
    cff = np.nonzero(is_f[ks] & (ks > 0))[0] + 1                 # rows calling a father, +1
    seg_start = np.concatenate([[2], cff + 1]).astype(np.int64)
    seg_end = np.concatenate([cff - 1, [N - 1]]).astype(np.int64)
    nrun = seg_start.size
    # Level of a run: 0 for leaf runs, else 1 + max level of the runs calling the
    # father the run starts at.
    level = np.zeros(nrun, dtype=np.int64)
    level[is_f[seg_start - 1]] = 1
    for i in range(1, 51):
        callers = np.nonzero(level == i)[0]
        if callers.size == 0:
            continue
        callees = ks[seg_end[callers]]
        mark = np.zeros(N + 2, dtype=bool)
        mark[callees] = True
        level[mark[seg_start - 1]] = i + 1
    # FLevel of a father: 1 + max level of the runs that call it (1 for leaf callers).
    rfather = {int(fa): k for k, fa in enumerate(fathers)}
    flevel = np.ones(fathers.size, dtype=np.int64)
    for i in range(1, int(level.max()) + 1 if nrun else 1):
        for j in np.nonzero(level == i)[0]:
            callee = int(ks[seg_end[j]])
            if callee in rfather:
                flevel[rfather[callee]] = i + 1
    seg_end[-1] = N
    depth = int(level.max())
    cbl, cbfl, lrel, flrel = [], [], [], []
    nrow = nfrow = 0
    for L in range(depth + 1):
        comps = list(np.nonzero(level == L)[0] + 1)
        if not comps:
            raise TreeError(f"level {L} has no runs")
        blk = _pad32(comps)
        cbl.append(blk)
        lrel.append(list(range(nrow + 1, nrow + blk.shape[0] + 1)))
        nrow += blk.shape[0]
        if L < depth:
            fcomps = list(np.nonzero(flevel == L + 1)[0] + 1)
            if not fcomps:
                raise TreeError(f"father level {L + 1} is empty")
            fblk = _pad32(fcomps)
            cbfl.append(fblk)
            flrel.append(list(range(nfrow + 1, nfrow + fblk.shape[0] + 1)))
            nfrow += fblk.shape[0]
    lrs, lre, _ = _cell_2_vec(lrel)
    if flrel:
        flrs, flre, _ = _cell_2_vec(flrel)
    else:
        flrs = flre = np.zeros(0, np.int64)
    last_run = int(np.nonzero(level == depth)[0][0])
    ksb = np.concatenate([ks, [seg_end[last_run]]]) - 1
    ksb = np.where(ksb < 0, 65535, ksb)                       # written as uint16
    return dict(
        fathers=fathers.astype(np.int64), rel_starts=rs + 1, rel_ends=re + 1, rel_vec=rv + 1,
        seg_start=seg_start, seg_end=seg_end, depth=depth, level=level, flevel=flevel,
        comp_by_level32=np.vstack(cbl) if cbl else np.zeros((0, 32), np.int64),
        comp_by_flevel32=np.vstack(cbfl) if cbfl else np.zeros((0, 32), np.int64),
        lrel_starts=lrs, lrel_ends=lre, flrel_starts=flrs, flrel_ends=flre, ksb_m1=ksb,
    )


def disperse_branching(par):
    """Give every node with more than two children (the root: more than one,
    since the root counts as its own first child) a chain of extension nodes
    that carries its children at most two per node, as the reference's
    model export does (``hinesDisperseBranching``, NeuroGPU_Base/runModel.hoc:158-215:
    an extension section of nchild - 1 segments at the node's end, child i
    reconnected to segment min(i, nchild - 2)), so that the solver's son
    numbering holds (create_auxilliary_data_3.py:72-84).

    Returns ``(new_par, src)`` in a new depth-first Hines order: ``src[i]``
    is the original segment of new node i, or ``-(h + 1)`` for an extension
    node of host segment h.  Trees that need no dispersal come back with the
    same order (``src = arange(N)``)."""
    par = np.asarray(par, dtype=np.int64)
    N = par.size
    if N < 1 or par[0] != -1 or np.any(par[1:] < 0) or np.any(par[1:] >= np.arange(1, N)):
        raise TreeError("parent vector must satisfy 0 <= parent(s) < s (Hines order) with the root first")
    kids = [[] for _ in range(N)]
    for c in range(1, N):
        kids[int(par[c])].append(c)
    # node ids: 0..N-1 original, N.. extension nodes; children lists in order
    ch = {i: list(kids[i]) for i in range(N)}
    host = {}
    nxt = N
    for p in range(N):
        k = len(kids[p])
        limit = 1 if p == 0 else 2
        if k <= limit:
            continue
        xs = list(range(nxt, nxt + k - 1))
        nxt += k - 1
        for x in xs:
            host[x] = p
        ch[p] = [xs[0]]
        for j, x in enumerate(xs):
            if j < len(xs) - 1:
                ch[x] = [kids[p][j], xs[j + 1]]
            else:
                ch[x] = [kids[p][j], kids[p][j + 1]]
    if not host:
        return par.copy(), np.arange(N, dtype=np.int64)
    order, newpar, stack = [], [], [(0, -1)]
    while stack:
        node, pnew = stack.pop()
        me = len(order)
        order.append(node)
        newpar.append(pnew)
        for c in reversed(ch[node]):
            stack.append((c, me))
    src = np.array([n if n < N else -(host[n] + 1) for n in order], dtype=np.int64)
    return np.array(newpar, dtype=np.int64), src


def build_tree(par, e, f, cm, seg_to_comp, bool_model, ncomps: int) -> TreeData:
    """Assemble a complete :class:`TreeData` for a morphology.

    ``par`` is the segment parent vector (root -1, parent(s) < s), ``e``/``f``
    the off-diagonal coefficients in Hines-row order, ``cm`` per segment,
    ``bool_model`` of shape [nmodels, N]."""
    par = np.asarray(par, dtype=np.int64)
    N = par.size
    ks = parent_to_ks(par)
    son = son_numbers(par)
    if son.max(initial=0) > 2:
        raise TreeError("a node has more than 2 children (or the root more than 1); "
                        "disperse the branching first (disperse_branching; create_auxilliary_data_3.py:83-84)")
    fidxs, logn = ancestor_table(ks)
    sch = schedule(ks)
    bm = np.asarray(bool_model, dtype=np.int64).reshape(-1, N)
    return TreeData(
        N=N, ncomps=int(ncomps), e=np.asarray(e, np.float64), f=np.asarray(f, np.float64),
        ks=ks, seg_to_comp=np.asarray(seg_to_comp, np.int64), cm=np.asarray(cm, np.float64),
        nmodels=bm.shape[0], bool_model=bm, son_no=son, depth=sch["depth"], logn_depth=logn,
        rel_starts=sch["rel_starts"], rel_ends=sch["rel_ends"], rel_vec=sch["rel_vec"],
        seg_start=sch["seg_start"], seg_end=sch["seg_end"], fathers=sch["fathers"],
        fidxs=fidxs, comp_by_level32=sch["comp_by_level32"],
        comp_by_flevel32=sch["comp_by_flevel32"], lrel_starts=sch["lrel_starts"],
        lrel_ends=sch["lrel_ends"], flrel_starts=sch["flrel_starts"],
        flrel_ends=sch["flrel_ends"], ksb_m1=sch["ksb_m1"],
    )


def rebuild_from(t: TreeData) -> TreeData:
    """Recompute every map of ``t`` from its parent vector alone."""
    return build_tree(t.parent_segments(), t.e, t.f, t.cm, t.seg_to_comp, t.bool_model, t.ncomps)
