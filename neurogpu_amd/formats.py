"""Readers and writers for NeuroGPU's data files (the file boundary of the
simulation loop).

Every format follows the reference byte for byte where it is read or written:

* ``BasicConstSegP.csv`` -- tree data, one record per line, values ``"%.9f,"``.
  Written by ``NeuroGPU_Base/scripts/pyscripts/create_auxilliary_data_3.py:229-272``,
  parsed by ``GUI/BBP_TTPC_EXAMPLE/pyNeuroGPU_unix/src/MainC.cu:335-551``.
* ``AllParams.csv`` -- ``"%5d,"`` NSets, then NSets rows of NComps*NParams values
  (index ``p*NComps + comp``); ``proc_add_param_to_hoc_for_opt.py:281-310``,
  parsed by ``src/Util.cu:210-234``.
* ``Sim.csv`` -- v_init, dt, nrecsites, recsites (``PYS/extractModel.py:674-692``,
  ``src/Util.cu:394-432``).
* ``Stim_meta.csv`` -- NStimuli, Nt, comp, loc (1-based), area (``src/Util.cu:337-357``).
* ``Stim_raw[<i>].csv`` -- NStimuli lines of at least Nt amplitudes (nA).
* ``times[<i>].csv`` -- one line of Nt step sizes (ms).
* ``VHotP<dev>.dat`` -- int32 N, int32 3, N float64, layout [set][stim][rec][t]
  (``src/Util.cu:443-476``).
* ``RunTimes.csv`` -- ``"NSets,ms\\n"`` (``src/CudaStuff.cu:574-581``).

Parsing is strict: a wrong value count raises ``FormatError`` instead of the
reference's silent truncation.
"""
from __future__ import annotations

import dataclasses
import os
import struct
from typing import List, Optional, Sequence

import numpy as np


class FormatError(ValueError):
    """A data file does not match the layout the simulator expects."""


def _split(line: str) -> List[str]:
    toks = [t.strip() for t in line.strip().split(",")]
    return [t for t in toks if t != ""]


def _ints(line: str, n: Optional[int] = None, what: str = "") -> np.ndarray:
    toks = _split(line)
    # The reference parses integers with atoi, i.e. it truncates "2.000000000" to 2.
    vals = np.array([int(float(t)) for t in toks], dtype=np.int64)
    if n is not None and vals.size != n:
        raise FormatError(f"{what}: expected {n} integers, found {vals.size}")
    return vals


def _floats(line: str, n: Optional[int] = None, what: str = "", at_least: bool = False) -> np.ndarray:
    vals = np.array([float(t) for t in _split(line)], dtype=np.float64)
    if n is not None:
        if at_least:
            if vals.size < n:
                raise FormatError(f"{what}: expected at least {n} values, found {vals.size}")
            vals = vals[:n]
        elif vals.size != n:
            raise FormatError(f"{what}: expected {n} values, found {vals.size}")
    return vals


@dataclasses.dataclass
class TreeData:
    """Contents of BasicConstSegP.csv.  Index conventions are the file's:
    ``e``, ``f``, ``ks`` and ``fidxs`` are in Hines-row order (row h = N-1-s for
    segment s), ``ks``/``fidxs``/``rel_*``/``seg_*``/``fathers``/``comp_by_*``
    hold 1-based row numbers; ``seg_to_comp``, ``cm``, ``bool_model`` and
    ``son_no`` are per segment."""

    N: int
    ncomps: int
    e: np.ndarray
    f: np.ndarray
    ks: np.ndarray
    seg_to_comp: np.ndarray
    cm: np.ndarray
    nmodels: int
    bool_model: np.ndarray          # [nmodels, N]
    son_no: np.ndarray
    depth: int
    logn_depth: int
    rel_starts: np.ndarray
    rel_ends: np.ndarray
    rel_vec: np.ndarray
    seg_start: np.ndarray
    seg_end: np.ndarray
    fathers: np.ndarray
    fidxs: np.ndarray               # [logn_depth, N]
    comp_by_level32: np.ndarray     # [comp_depth, 32]
    comp_by_flevel32: np.ndarray    # [comp_fdepth, 32]
    lrel_starts: np.ndarray
    lrel_ends: np.ndarray
    flrel_starts: np.ndarray
    flrel_ends: np.ndarray
    ksb_m1: np.ndarray

    @property
    def n_fathers(self) -> int:
        return int(self.fathers.size)

    @property
    def n_call(self) -> int:
        return int(self.rel_vec.size)

    def parent_segments(self) -> np.ndarray:
        """Parent of each segment (segment space, root = -1):
        ``parent(s) = N - Ks[N - s]`` (``src/CudaStuff.cu:195``)."""
        N = self.N
        par = np.full(N, -1, dtype=np.int64)
        for s in range(1, N):
            par[s] = N - int(self.ks[N - s])
        return par


def read_tree(path: str) -> TreeData:
    """Parse ``BasicConstSegP.csv`` (``src/MainC.cu:335-551``)."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    if len(lines) < 31:
        raise FormatError(f"{path}: expected at least 31 lines, found {len(lines)}")
    it = iter(lines)
    nx = lambda: next(it)
    N = int(_ints(nx(), 1, "N")[0])
    ncomps = int(_ints(nx(), 1, "NComps")[0])
    if N < 1 or ncomps < 1:
        raise FormatError("N and NComps must be positive")
    e = _floats(nx(), N, "e")
    f = _floats(nx(), N, "f")
    ks = _ints(nx(), N, "Ks")
    seg_to_comp = _ints(nx(), N, "SegToComp")
    cm = _floats(nx(), N, "Cms")
    nmodels = int(_ints(nx(), 1, "NModels")[0])
    bool_model = _ints(nx(), nmodels * N, "boolModel").reshape(nmodels, N)
    son_no = _ints(nx(), N, "SonNoVec")
    depth = int(_ints(nx(), 1, "Depth")[0])
    logn = int(_ints(nx(), 1, "LognDepth")[0])
    nf = int(_ints(nx(), 1, "nFathers")[0])
    ncall = int(_ints(nx(), 1, "nCallForFather")[0])
    rel_starts = _ints(nx(), nf, "RelStarts")
    rel_ends = _ints(nx(), nf, "RelEnds")
    rel_vec = _ints(nx(), ncall, "RelVec")
    seg_start = _ints(nx(), ncall + 1, "SegStartI")
    seg_end = _ints(nx(), ncall + 1, "SegEndI")
    fathers = _ints(nx(), nf, "Fathers")
    fidxs = _ints(nx(), logn * N, "FIdxs").reshape(logn, N)
    cdepth = int(_ints(nx(), 1, "CompDepth")[0])
    cfdepth = int(_ints(nx(), 1, "CompFDepth")[0])
    cbl = _ints(nx(), cdepth * 32, "CompByLevel32").reshape(cdepth, 32)
    cbfl = _ints(nx(), cfdepth * 32, "CompByFLevel32").reshape(cfdepth, 32)
    nl = int(_ints(nx(), 1, "nLRel")[0])
    lrs = _ints(nx(), nl, "LRelStarts")
    lre = _ints(nx(), nl, "LRelEnds")
    nfl = int(_ints(nx(), 1, "nFLRel")[0])
    flrs = _ints(nx(), nfl, "FLRelStarts")
    flre = _ints(nx(), nfl, "FLRelEnds")
    try:
        ksb = _ints(nx())
    except StopIteration:
        ksb = np.zeros(0, dtype=np.int64)
    return TreeData(N, ncomps, e, f, ks, seg_to_comp, cm, nmodels, bool_model, son_no,
                    depth, logn, rel_starts, rel_ends, rel_vec, seg_start, seg_end, fathers,
                    fidxs, cbl, cbfl, lrs, lre, flrs, flre, ksb)


def _fmt9(vals: Sequence[float]) -> str:
    # numpy.savetxt(fmt='%.9f', newline=',') of a flattened array, as the generator writes it.
    return "".join("%.9f," % float(v) for v in np.asarray(vals).ravel())


def write_tree(path: str, t: TreeData) -> None:
    """Write ``BasicConstSegP.csv`` in the generator's layout
    (``create_auxilliary_data_3.py:229-272``)."""
    rows = [
        [t.N], [t.ncomps], t.e, t.f, t.ks, t.seg_to_comp, t.cm, [t.nmodels],
        np.asarray(t.bool_model).ravel(), t.son_no, [t.depth], [t.logn_depth],
        [t.n_fathers], [t.n_call], t.rel_starts, t.rel_ends, t.rel_vec, t.seg_start,
        t.seg_end, t.fathers, np.asarray(t.fidxs).ravel(), [t.comp_by_level32.shape[0]],
        [t.comp_by_flevel32.shape[0]], t.comp_by_level32.ravel(), t.comp_by_flevel32.ravel(),
        [t.lrel_starts.size], t.lrel_starts, t.lrel_ends, [t.flrel_starts.size],
        t.flrel_starts, t.flrel_ends, t.ksb_m1,
    ]
    with open(path, "w") as fh:
        fh.write("".join(_fmt9(r) + "\n" for r in rows))


# ---------------------------------------------------------------- AllParams

def read_allparams(path: str, ncomps: int, nparams: int) -> np.ndarray:
    """Parse ``AllParams.csv`` into ``[nsets, nparams*ncomps]`` (``src/Util.cu:210-234``)."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    if not lines:
        raise FormatError(f"{path}: empty file")
    nsets = int(_ints(lines[0], 1, "NSets")[0])
    if nsets < 0 or len(lines) - 1 < nsets:
        raise FormatError(f"{path}: NSets={nsets} but {len(lines) - 1} parameter rows")
    out = np.empty((nsets, nparams * ncomps), dtype=np.float64)
    for i in range(nsets):
        out[i] = _floats(lines[1 + i], nparams * ncomps, f"AllParams row {i}")
    return out


def write_allparams(path: str, params: np.ndarray) -> None:
    """Write ``AllParams.csv``: ``"%5d,"`` header, rows of ``"%.5e,"`` values
    (``proc_add_param_to_hoc_for_opt.py:302-310``).  Values are quantised to
    6 significant digits, exactly as the reference stores them."""
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    with open(path, "w") as fh:
        fh.write("%5d,\n" % params.shape[0])
        for row in params:
            fh.write("".join("%.5e," % v for v in row) + "\n")


def quantise_params(params: np.ndarray) -> np.ndarray:
    """The values a round trip through AllParams.csv produces (``%.5e``)."""
    flat = np.asarray(params, dtype=np.float64).ravel()
    return np.array([float("%.5e" % v) for v in flat]).reshape(np.shape(params))


def read_allparams_dat(path: str, ncomps: int, nparams: int) -> np.ndarray:
    """``AllParams.dat``: uint16 NSets then float32 values (``src/Util.cu:317-320``)."""
    raw = open(path, "rb").read()
    if len(raw) < 2:
        raise FormatError(f"{path}: too short")
    nsets = struct.unpack("<H", raw[:2])[0]
    need = 2 + 4 * nsets * ncomps * nparams
    if len(raw) != need:
        raise FormatError(f"{path}: {len(raw)} bytes, expected {need}")
    return np.frombuffer(raw[2:], dtype="<f4").astype(np.float64).reshape(nsets, nparams * ncomps)


# ---------------------------------------------------------------- Sim / Stim

@dataclasses.dataclass
class SimData:
    v_init: float
    dt: float
    rec_sites: np.ndarray


def read_sim(path: str) -> SimData:
    """``Sim.csv`` (``src/Util.cu:394-432``).  The reference reads record sites
    through a binary fread of the text; the build reads the text values."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    if len(lines) < 3:
        raise FormatError(f"{path}: expected at least 3 lines")
    v_init = float(_floats(lines[0], 1, "v_init")[0])
    dt = float(_floats(lines[1], 1, "dt")[0])
    nrec = int(_ints(lines[2], 1, "nrecsite")[0])
    rec = _ints(lines[3], nrec, "recsites") if nrec > 0 else np.zeros(0, np.int64)
    return SimData(v_init, dt, rec)


def write_sim(path: str, sim: SimData) -> None:
    with open(path, "w") as fh:
        fh.write("%.5f,\n%.5f,\n%.5f,\n" % (sim.v_init, sim.dt, len(sim.rec_sites)))
        fh.write("".join("%.5f," % r for r in sim.rec_sites) + "\n")


@dataclasses.dataclass
class StimData:
    nstim: int
    nt: int
    comp: int
    loc: int            # 1-based segment, as in Stim_meta.csv
    area: float
    amps: np.ndarray    # [nstim, nt]
    dts: np.ndarray     # [nt]

    @property
    def stim_seg(self) -> int:
        """0-based segment that receives the current (``src/Util.cu:355``)."""
        return self.loc - 1


def read_stim(data_dir: str, suffix: str = "") -> StimData:
    """``Stim_meta.csv`` + ``Stim_raw<suffix>.csv`` + ``times<suffix>.csv``
    (``src/Util.cu:337-390``).  Rows of Stim_raw may hold more than Nt values;
    the extras are ignored as in the reference."""
    with open(os.path.join(data_dir, "Stim_meta.csv")) as fh:
        meta = fh.read().splitlines()
    if len(meta) < 5:
        raise FormatError("Stim_meta.csv: expected 5 lines")
    nstim = int(_ints(meta[0], 1, "NStimuli")[0])
    nt = int(float(_floats(meta[1], 1, "Nt")[0]))
    comp = int(_ints(meta[2], 1, "comp")[0])
    loc = int(_ints(meta[3], 1, "loc")[0])
    area = float(_floats(meta[4], 1, "area")[0])
    if nstim < 1 or nt < 1:
        raise FormatError("Stim_meta.csv: NStimuli and Nt must be positive")
    if not area > 0:
        raise FormatError("Stim_meta.csv: stimulus area must be positive")
    with open(os.path.join(data_dir, f"Stim_raw{suffix}.csv")) as fh:
        raw = fh.read().splitlines()
    if len(raw) < nstim:
        raise FormatError(f"Stim_raw{suffix}.csv: expected {nstim} rows")
    amps = np.stack([_floats(raw[i], nt, f"Stim_raw row {i}", at_least=True) for i in range(nstim)])
    tname = os.path.join(data_dir, f"times{suffix}.csv")
    if not os.path.exists(tname):
        tname = os.path.join(data_dir, "times.csv")
    with open(tname) as fh:
        dts = _floats(fh.readline(), nt, "times", at_least=True)
    if np.any(~(dts > 0)):
        raise FormatError("times.csv: every dt must be positive")
    return StimData(nstim, nt, comp, loc, area, amps, dts)


def write_stim(data_dir: str, stim: StimData, suffix: str = "") -> None:
    with open(os.path.join(data_dir, "Stim_meta.csv"), "w") as fh:
        fh.write(f"{stim.nstim}\n{stim.nt}\n{stim.comp}\n{stim.loc}\n{stim.area!r}\n")
    with open(os.path.join(data_dir, f"Stim_raw{suffix}.csv"), "w") as fh:
        for row in np.atleast_2d(stim.amps):
            fh.write("".join("%.6f," % a for a in row) + "\n")
    with open(os.path.join(data_dir, f"times{suffix}.csv"), "w") as fh:
        fh.write("".join("%.5f," % d for d in stim.dts) + "\n")


# ---------------------------------------------------------------- outputs

def write_vhot(path: str, vhot: np.ndarray) -> None:
    """``VHotP<dev>.dat``: int32 N, int32 3, N doubles (``src/Util.cu:443-456``)."""
    arr = np.ascontiguousarray(vhot, dtype="<f8").ravel()
    with open(path, "wb") as fh:
        fh.write(struct.pack("<ii", arr.size, 3))
        fh.write(arr.tobytes())


def read_vhot(path: str) -> np.ndarray:
    """Inverse of :func:`write_vhot` (the evaluators' ``nrnMread``)."""
    raw = open(path, "rb").read()
    if len(raw) < 8:
        raise FormatError(f"{path}: too short")
    n, prec = struct.unpack("<ii", raw[:8])
    if len(raw) != 8 + 8 * n:
        raise FormatError(f"{path}: header says {n} values, file holds {(len(raw) - 8) / 8}")
    return np.frombuffer(raw[8:], dtype="<f8").copy()


def write_runtimes(path: str, nsets: int, ms: float) -> None:
    with open(path, "w") as fh:
        fh.write(f"{nsets},{ms}\n")
