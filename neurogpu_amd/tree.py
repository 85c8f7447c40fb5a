"""Tree maps of a morphology: from a segment parent vector to the
Hines-ordering maps of ``BasicConstSegP.csv``, computed by the library's C++
scheduler (``neurogpu_amd/csrc/treemap.h`` behind ``ngpu_tree_build``,
``ngpu_tree_disperse`` and ``ngpu_tree_write``).

The maps are the reference generator's
(``NeuroGPU_Base/scripts/pyscripts/create_auxilliary_data_3.py:22-315`` and
``make_tree_from_parent_vec.py:16-59``), so that a morphology built here can
be written in the reference's file format, read by the reference's parser and
solved by the reference's BeforeLU/BkSub schedule.  Wide branch points are
dispersed as the reference's model export does
(``NeuroGPU_Base/runModel.hoc:158-215``).

Conventions (all as in the file): row h = N-1-s for segment s; ``Ks[r]`` is the
1-based parent row of 1-based row r (r = 1..N-1; ``Ks[0]`` = 0); run and father
tables are 1-based.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .formats import TreeData


class TreeError(ValueError):
    """The parent vector cannot be scheduled by the reference's solver."""


def parent_to_ks(par: np.ndarray) -> np.ndarray:
    """Ks of a segment parent vector (root -1): ``Ks[j] = N - parent(N - j)``,
    the inverse of ``parent(s) = N - Ks[N - s]`` (create_auxilliary_data_3.py:45-50)."""
    par = np.asarray(par, dtype=np.int64)
    N = par.size
    if N < 1 or par[0] != -1 or np.any(par[1:] < 0) or np.any(par[1:] >= np.arange(1, N)):
        raise TreeError("parent vector must satisfy 0 <= parent(s) < s (Hines order) with the root first")
    ks = np.zeros(N, dtype=np.int64)
    ks[1:] = N - par[N - np.arange(1, N)]
    return ks


def disperse_branching(par):
    """``(new_par, src)``: every node with more than two children (the root:
    more than one, since the root counts as its own first child) gets a chain
    of extension nodes that carries its children at most two per node, in a
    new depth-first Hines order; ``src[i]`` is the original segment of new node
    i, or ``-(h + 1)`` for an extension node of host segment h.  Trees that
    need no dispersal come back unchanged (``src = arange(N)``)."""
    try:
        return _lib.tree_disperse(par)
    except _lib.NgpuError as ex:
        raise TreeError(str(ex)) from None


def build_tree(par, e, f, cm, seg_to_comp, bool_model, ncomps: int) -> TreeData:
    """A complete :class:`TreeData` for a morphology: ``par`` is the segment
    parent vector (root -1, parent(s) < s, at most two children a node and one
    on the root: :func:`disperse_branching` first otherwise), ``e``/``f`` the
    off-diagonal coefficients in Hines-row order, ``cm`` per segment,
    ``bool_model`` of shape [nmodels, N]."""
    try:
        return _lib.tree_build(par, e, f, cm, seg_to_comp, bool_model, ncomps)
    except _lib.NgpuError as ex:
        raise TreeError(str(ex)) from None


def write_tree(path: str, t: TreeData) -> None:
    """``BasicConstSegP.csv`` of a tree, by the library's writer."""
    try:
        _lib.tree_write(t, path)
    except _lib.NgpuError as ex:
        raise TreeError(str(ex)) from None


def rebuild_from(t: TreeData) -> TreeData:
    """Recompute every map of ``t`` from its parent vector alone."""
    return build_tree(t.parent_segments(), t.e, t.f, t.cm, t.seg_to_comp, t.bool_model, t.ncomps)
