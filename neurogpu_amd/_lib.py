"""ctypes binding of libneurogpu (the C ABI declared in include/neurogpu.h).

The library is built in-tree (``neurogpu_amd/_build/libneurogpu.so``) by
:func:`neurogpu_amd.build.build`.  There is no fallback: if the library is
missing or cannot create a context, these calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .formats import TreeData

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NGPU_LIB") or os.path.join(_HERE, "_build", "libneurogpu.so")

_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

EXPORTS = ("ngpu_create", "ngpu_set_tree", "ngpu_set_sim", "ngpu_set_params", "ngpu_set_stimuli",
           "ngpu_set_launch", "ngpu_run", "ngpu_fetch_vhot", "ngpu_open", "ngpu_write_vhot",
           "ngpu_vhot_size", "ngpu_mechset_nparams", "ngpu_last_error", "ngpu_close", "ngpu_features",
           "ngpu_mlp_create", "ngpu_mlp_nparams", "ngpu_mlp_set_weights", "ngpu_mlp_get_weights", "ngpu_mlp_train",
           "ngpu_mlp_predict", "ngpu_mlp_gradients", "ngpu_mlp_last_error", "ngpu_mlp_close", "ngpu_gemm_bf16",
           "ngpu_batch_create", "ngpu_batch_add", "ngpu_batch_set_launch", "ngpu_batch_run",
           "ngpu_batch_member_info", "ngpu_batch_last_error", "ngpu_batch_close", "ngpu_run_steps",
           "ngpu_get_sizes", "ngpu_features_noalloc", "ngpu_set_solve",
           "ngpu_set_param_template", "ngpu_set_genomes", "ngpu_tree_build", "ngpu_tree_free",
           "ngpu_tree_write", "ngpu_tree_disperse")
NFEAT = 5
FEATURES = ("spike_count", "first_latency", "ap_amplitude", "ap_width", "mean_isi")


class NgpuError(RuntimeError):
    """A libneurogpu call returned an error code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libneurogpu error {code}: {msg}")
        self.code = code


class ngpu_tree(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ncomps", C.c_int), ("mechset", C.c_int), ("nmodels", C.c_int),
        ("e", _DP), ("f", _DP), ("cm", _DP),
        ("ks", _IP), ("seg_to_comp", _IP), ("bool_model", _IP), ("son_no", _IP),
        ("depth", C.c_int), ("logn_depth", C.c_int), ("n_fathers", C.c_int), ("n_call", C.c_int),
        ("rel_starts", _IP), ("rel_ends", _IP), ("rel_vec", _IP), ("seg_start", _IP),
        ("seg_end", _IP), ("fathers", _IP), ("fidxs", _IP),
        ("comp_depth", C.c_int), ("comp_fdepth", C.c_int),
        ("comp_by_level32", _IP), ("comp_by_flevel32", _IP),
        ("n_lrel", C.c_int), ("n_flrel", C.c_int),
        ("lrel_starts", _IP), ("lrel_ends", _IP), ("flrel_starts", _IP), ("flrel_ends", _IP),
    ]


class ngpu_stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_kernel", C.c_double), ("comp_steps", C.c_double),
                ("comp_steps_per_s", C.c_double), ("waves_per_run", C.c_int), ("slots_per_lane", C.c_int),
                ("runs_per_block", C.c_int), ("launches", C.c_int), ("nvar", C.c_int),
                ("nsets", C.c_int), ("nstim", C.c_int), ("nrec", C.c_int), ("nt", C.c_int), ("N", C.c_int),
                ("runs_per_group", C.c_int), ("solve", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


SOLVE_AUTO, SOLVE_REFERENCE, SOLVE_FAST = 0, 1, 2
MECH_PASSIVE, MECH_TTPC, MECH_MAINEN, MECH_TTPC_GEN, MECH_MAINEN_GEN = 0, 1, 2, 3, 4


class ngpu_batch_launch(C.Structure):
    _fields_ = [("waves_per_run", C.c_int), ("block_sync", C.c_int), ("lds_staging", C.c_int),
                ("runs_per_group", C.c_int)]


class ngpu_batch_stats(C.Structure):
    _fields_ = [("ms_kernel", C.c_double), ("comp_steps", C.c_double), ("nitems", C.c_int),
                ("workgroups", C.c_int), ("lds_bytes", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ngpu_batch_member(C.Structure):
    _fields_ = [("first_start_ms", C.c_double), ("last_end_ms", C.c_double), ("comp_steps", C.c_double),
                ("waves_per_run", C.c_int), ("runs_per_item", C.c_int), ("block_sync", C.c_int),
                ("lds_staging", C.c_int), ("runs_per_group", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libneurogpu is not built ({LIB_PATH}); run neurogpu_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ngpu_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.ngpu_set_tree.argtypes = [vp, C.POINTER(ngpu_tree)]
        L.ngpu_set_sim.argtypes = [vp, C.c_double, C.c_int, _IP]
        L.ngpu_set_params.argtypes = [vp, _DP, C.c_int, C.c_int, C.c_int]
        L.ngpu_set_stimuli.argtypes = [vp, _DP, C.c_int, C.c_int, _DP, C.c_int, C.c_int, C.c_double]
        L.ngpu_set_launch.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ngpu_set_solve.argtypes = [vp, C.c_int]
        L.ngpu_set_param_template.argtypes = [vp, _DP, C.c_int, C.c_int, C.c_int]
        L.ngpu_tree_build.argtypes = [C.c_int, _IP, _DP, _DP, _DP, _IP, C.c_int, C.c_int, _IP,
                                      C.POINTER(C.POINTER(ngpu_tree))]
        L.ngpu_tree_free.argtypes = [C.POINTER(ngpu_tree)]
        L.ngpu_tree_free.restype = None
        L.ngpu_tree_write.argtypes = [C.POINTER(ngpu_tree), C.c_char_p]
        L.ngpu_tree_disperse.argtypes = [C.c_int, _IP, C.c_int, _IP, _IP, C.POINTER(C.c_int)]
        L.ngpu_set_genomes.argtypes = [vp, _DP, C.c_int]
        L.ngpu_run.argtypes = [vp, _DP, C.POINTER(ngpu_stats)]
        L.ngpu_run_steps.argtypes = [vp, C.c_int, C.c_int, C.POINTER(ngpu_stats)]
        L.ngpu_get_sizes.argtypes = [vp] + [_IP] * 5
        L.ngpu_fetch_vhot.argtypes = [vp, _DP]
        L.ngpu_open.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(vp)]
        L.ngpu_write_vhot.argtypes = [vp, C.c_char_p]
        L.ngpu_vhot_size.argtypes = [vp]
        L.ngpu_vhot_size.restype = C.c_longlong
        L.ngpu_mechset_nparams.argtypes = [C.c_int]
        L.ngpu_last_error.argtypes = [vp]
        L.ngpu_last_error.restype = C.c_char_p
        L.ngpu_features.argtypes = [vp, _DP, _DP, C.c_double, _DP]
        L.ngpu_features.restype = C.c_int
        L.ngpu_features_noalloc.argtypes = [vp, _DP, _DP, C.c_double, _DP]
        L.ngpu_features_noalloc.restype = C.c_int
        L.ngpu_close.argtypes = [vp]
        L.ngpu_close.restype = None
        L.ngpu_batch_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.ngpu_batch_add.argtypes = [vp, vp]
        L.ngpu_batch_set_launch.argtypes = [vp, C.c_int, C.POINTER(ngpu_batch_launch)]
        L.ngpu_batch_run.argtypes = [vp, C.POINTER(ngpu_batch_stats)]
        L.ngpu_batch_member_info.argtypes = [vp, C.c_int, C.POINTER(ngpu_batch_member)]
        L.ngpu_batch_last_error.argtypes = [vp]
        L.ngpu_batch_last_error.restype = C.c_char_p
        L.ngpu_batch_close.argtypes = [vp]
        L.ngpu_batch_close.restype = None
        for name in ("ngpu_batch_create", "ngpu_batch_add", "ngpu_batch_set_launch", "ngpu_batch_run",
                     "ngpu_batch_member_info"):
            getattr(L, name).restype = C.c_int
        for name in ("ngpu_create", "ngpu_set_tree", "ngpu_set_sim", "ngpu_set_params", "ngpu_set_stimuli",
                     "ngpu_set_launch", "ngpu_run", "ngpu_fetch_vhot", "ngpu_open", "ngpu_write_vhot",
                     "ngpu_mechset_nparams", "ngpu_run_steps", "ngpu_get_sizes", "ngpu_set_solve",
                     "ngpu_set_param_template", "ngpu_set_genomes", "ngpu_tree_build", "ngpu_tree_free",
           "ngpu_tree_write", "ngpu_tree_disperse"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _tree_struct(t: TreeData, mechset: int):
    keep = []

    def ia(x):
        a = np.ascontiguousarray(np.asarray(x).ravel(), dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(_IP)

    def da(x):
        a = np.ascontiguousarray(np.asarray(x).ravel(), dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(_DP)

    s = ngpu_tree(t.N, t.ncomps, mechset, t.nmodels, da(t.e), da(t.f), da(t.cm), ia(t.ks), ia(t.seg_to_comp),
                  ia(t.bool_model), ia(t.son_no), t.depth, t.logn_depth, t.n_fathers, t.n_call,
                  ia(t.rel_starts), ia(t.rel_ends), ia(t.rel_vec), ia(t.seg_start), ia(t.seg_end),
                  ia(t.fathers), ia(t.fidxs), t.comp_by_level32.shape[0], t.comp_by_flevel32.shape[0],
                  ia(t.comp_by_level32), ia(t.comp_by_flevel32), t.lrel_starts.size, t.flrel_starts.size,
                  ia(t.lrel_starts), ia(t.lrel_ends), ia(t.flrel_starts), ia(t.flrel_ends))
    return s, keep


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x).ravel(), dtype=np.int32)


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x).ravel(), dtype=np.float64)


def tree_build(par, e, f, cm, seg_to_comp, bool_model, ncomps: int) -> TreeData:
    """Every map of BasicConstSegP.csv for a morphology, computed by the
    library's tree-map scheduler (ngpu_tree_build)."""
    L = lib()
    par, e, f, cm, s2c = _i32(par), _f64(e), _f64(f), _f64(cm), _i32(seg_to_comp)
    N = par.size
    bm = np.ascontiguousarray(np.asarray(bool_model).reshape(-1, N), dtype=np.int32)
    out = C.POINTER(ngpu_tree)()
    rc = L.ngpu_tree_build(N, par.ctypes.data_as(_IP), e.ctypes.data_as(_DP), f.ctypes.data_as(_DP),
                           cm.ctypes.data_as(_DP), s2c.ctypes.data_as(_IP), int(ncomps), bm.shape[0],
                           bm.ctypes.data_as(_IP), C.byref(out))
    if rc != 0:
        raise NgpuError(rc, (L.ngpu_last_error(None) or b"").decode())
    try:
        t = out.contents

        def ints(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int64) if n else np.zeros(0, np.int64)

        nl, nfl = t.n_lrel, t.n_flrel
        td = TreeData(
            N=t.N, ncomps=t.ncomps, e=e.copy(), f=f.copy(), ks=ints(t.ks, N), seg_to_comp=ints(t.seg_to_comp, N),
            cm=cm.copy(), nmodels=t.nmodels, bool_model=ints(t.bool_model, t.nmodels * N).reshape(t.nmodels, N),
            son_no=ints(t.son_no, N), depth=t.depth, logn_depth=t.logn_depth,
            rel_starts=ints(t.rel_starts, t.n_fathers), rel_ends=ints(t.rel_ends, t.n_fathers),
            rel_vec=ints(t.rel_vec, t.n_call), seg_start=ints(t.seg_start, t.n_call + 1),
            seg_end=ints(t.seg_end, t.n_call + 1), fathers=ints(t.fathers, t.n_fathers),
            fidxs=ints(t.fidxs, t.logn_depth * N).reshape(t.logn_depth, N),
            comp_by_level32=ints(t.comp_by_level32, t.comp_depth * 32).reshape(t.comp_depth, 32),
            comp_by_flevel32=ints(t.comp_by_flevel32, t.comp_fdepth * 32).reshape(t.comp_fdepth, 32),
            lrel_starts=ints(t.lrel_starts, nl), lrel_ends=ints(t.lrel_ends, nl),
            flrel_starts=ints(t.flrel_starts, nfl), flrel_ends=ints(t.flrel_ends, nfl), ksb_m1=np.zeros(0, np.int64))
    finally:
        L.ngpu_tree_free(out)
    # the file's last record (Ks - 1 as uint16, then the deepest level's first run end - 1)
    last_run = int(td.comp_by_level32[td.lrel_starts[td.depth], 0]) - 1
    ksb = np.concatenate([td.ks, [td.seg_end[last_run]]]) - 1
    td.ksb_m1 = np.where(ksb < 0, 65535, ksb)
    return td


def tree_write(t: TreeData, path: str) -> None:
    """Write BasicConstSegP.csv with the library's writer (ngpu_tree_write)."""
    L = lib()
    s, keep = _tree_struct(t, 0)
    rc = L.ngpu_tree_write(C.byref(s), path.encode())
    del keep
    if rc != 0:
        raise NgpuError(rc, (L.ngpu_last_error(None) or b"").decode())


def tree_disperse(par):
    """(new_par, src) of ngpu_tree_disperse."""
    L = lib()
    par = _i32(par)
    cap = 2 * max(par.size, 1)
    newp, src, n = np.zeros(cap, np.int32), np.zeros(cap, np.int32), C.c_int(0)
    rc = L.ngpu_tree_disperse(par.size, par.ctypes.data_as(_IP), cap, newp.ctypes.data_as(_IP),
                              src.ctypes.data_as(_IP), C.byref(n))
    if rc != 0:
        raise NgpuError(rc, (L.ngpu_last_error(None) or b"").decode())
    return newp[:n.value].astype(np.int64), src[:n.value].astype(np.int64)


class Simulator:
    """One libneurogpu context on one device.

    Mirrors the reference's run: tree data, Sim data, parameter matrix and
    stimuli in, VHot traces [nsets, nstim, nrec, nt] out."""

    def __init__(self, device: int = 0, _handle: Optional[int] = None):
        self._L = lib()
        self._h = C.c_void_p(_handle)
        if _handle is None:
            self._check(self._L.ngpu_create(int(device), C.byref(self._h)), None)
        self.nsets = self.nstim = self.nrec = self.nt = self.N = 0
        if _handle is not None:
            self._sync_sizes()

    def _sync_sizes(self) -> None:
        """Read the sizes the context holds (after ngpu_open the files set them)."""
        v = [C.c_int() for _ in range(5)]
        self._check(self._L.ngpu_get_sizes(self._h, *[C.byref(x) for x in v]))
        self.nsets, self.nstim, self.nrec, self.nt, self.N = (x.value for x in v)

    def _check(self, rc: int, h=...):
        if rc != 0:
            hh = self._h if h is ... else h
            msg = self._L.ngpu_last_error(hh)
            raise NgpuError(rc, msg.decode() if msg else "")

    @classmethod
    def open(cls, data_dir: str, suffix: str = "", mechset: int = -1, device: int = 0) -> "Simulator":
        """Load a reference data directory (ngpu_open)."""
        L = lib()
        h = C.c_void_p()
        rc = L.ngpu_open(data_dir.encode(), suffix.encode(), int(mechset), int(device), C.byref(h))
        if rc != 0:
            raise NgpuError(rc, (L.ngpu_last_error(None) or b"").decode())
        return cls(device, _handle=h.value)

    def set_tree(self, t: TreeData, mechset: int) -> None:
        s, keep = _tree_struct(t, mechset)
        self._check(self._L.ngpu_set_tree(self._h, C.byref(s)))
        del keep

    def set_sim(self, v_init: float, rec_sites) -> None:
        rec = np.ascontiguousarray(rec_sites, dtype=np.int32)
        self._check(self._L.ngpu_set_sim(self._h, float(v_init), rec.size, rec.ctypes.data_as(_IP)))
        self.nrec = rec.size

    def set_params(self, params: np.ndarray, ncomps: int, nparams: int) -> None:
        P = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        self._check(self._L.ngpu_set_params(self._h, P.ctypes.data_as(_DP), P.shape[0], int(ncomps), int(nparams)))
        self.nsets = P.shape[0]

    def set_param_template(self, template: np.ndarray, ncomps: int, nparams: int, nfree: int) -> None:
        """Template parameter row with free parameter i at the entries holding
        12300 + i (ngpu_set_param_template)."""
        T = np.ascontiguousarray(np.ravel(template), dtype=np.float64)
        self._check(self._L.ngpu_set_param_template(self._h, T.ctypes.data_as(_DP), int(ncomps), int(nparams),
                                                    int(nfree)))

    def set_genomes(self, x: np.ndarray) -> None:
        """Free parameters of each set [nsets, nfree]; assembled into the
        parameter table on the device (ngpu_set_genomes)."""
        X = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        self._check(self._L.ngpu_set_genomes(self._h, X.ctypes.data_as(_DP), X.shape[0]))
        self.nsets = X.shape[0]

    def set_stimuli(self, amps: np.ndarray, dts: np.ndarray, stim_seg: int, area: float) -> None:
        A = np.ascontiguousarray(np.atleast_2d(amps), dtype=np.float64)
        D = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(self._L.ngpu_set_stimuli(self._h, A.ctypes.data_as(_DP), A.shape[0], A.shape[1],
                                             D.ctypes.data_as(_DP), D.size, int(stim_seg), float(area)))
        self.nstim, self.nt = A.shape[0], D.size

    def set_launch(self, waves_per_run: int = 0, runs_per_group: int = 0, steps_per_launch: int = 0,
                   use_graph: bool = False) -> None:
        self._check(self._L.ngpu_set_launch(self._h, int(waves_per_run), int(runs_per_group),
                                            int(steps_per_launch), int(bool(use_graph))))

    def set_solve(self, mode: int) -> None:
        """Form of the Hines elimination: SOLVE_AUTO, SOLVE_REFERENCE or SOLVE_FAST."""
        self._check(self._L.ngpu_set_solve(self._h, int(mode)))

    def vhot_size(self) -> int:
        return int(self._L.ngpu_vhot_size(self._h))

    def run(self, fetch: bool = True):
        """Simulate; returns (vhot or None, stats dict)."""
        st = ngpu_stats()
        out = None
        if fetch:
            out = np.empty(self.vhot_size(), dtype=np.float64)
            self._check(self._L.ngpu_run(self._h, out.ctypes.data_as(_DP), C.byref(st)))
        else:
            self._check(self._L.ngpu_run(self._h, None, C.byref(st)))
        return out, st.as_dict()

    def run_steps(self, t_begin: int, t_end: int) -> dict:
        """Advance every run through steps [t_begin, t_end) (ngpu_run_steps);
        t_begin = 0 starts from the initial state, otherwise the state the
        previous window left at t_begin is continued.  Returns the stats."""
        st = ngpu_stats()
        self._check(self._L.ngpu_run_steps(self._h, int(t_begin), int(t_end), C.byref(st)))
        return st.as_dict()

    def fetch(self) -> np.ndarray:
        out = np.empty(self.vhot_size(), dtype=np.float64)
        self._check(self._L.ngpu_fetch_vhot(self._h, out.ctypes.data_as(_DP)))
        return out

    def features(self, t_on, t_off, thr: float = -30.0) -> np.ndarray:
        """Spike features of the last run's traces, computed on the device:
        [nsets, nstim, 5] (see FEATURES); spikes count in [t_on, t_off)."""
        self._sync_sizes()
        on = np.ascontiguousarray(np.broadcast_to(np.asarray(t_on, np.float64), (self.nstim,)))
        off = np.ascontiguousarray(np.broadcast_to(np.asarray(t_off, np.float64), (self.nstim,)))
        out = np.empty((self.nsets, self.nstim, NFEAT), dtype=np.float64)
        self._check(self._L.ngpu_features(self._h, on.ctypes.data_as(_DP), off.ctypes.data_as(_DP), float(thr),
                                          out.ctypes.data_as(_DP)))
        return out

    def write_vhot(self, path: str) -> None:
        self._check(self._L.ngpu_write_vhot(self._h, path.encode()))

    def close(self) -> None:
        if self._h and self._h.value:
            self._L.ngpu_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate_workload(w, device: int = 0, launch: Optional[dict] = None, fetch: bool = True,
                      solve: int = SOLVE_AUTO, mechset: Optional[int] = None):
    """Run a :class:`models.Workload` on the GPU; returns (vhot [nsets,nstim,nrec,nt], stats).
    ``mechset`` overrides the model's mechanism set id with one of the same
    mechanisms and parameter layout (MECH_TTPC_GEN for a TTPC model)."""
    m = w.model
    with Simulator(device) as sim:
        sim.set_tree(m.tree, m.mechset_id if mechset is None else int(mechset))
        sim.set_sim(m.v_init, m.rec_sites)
        sim.set_params(w.params, m.ncomps, m.nparams)
        sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
        sim.set_solve(solve)
        if launch:
            sim.set_launch(**launch)
        out, st = sim.run(fetch=fetch)
    if out is not None:
        out = out.reshape(w.x.shape[0], w.amps.shape[0], len(m.rec_sites), w.dts.size)
    return out, st


class Batch:
    """Several :class:`Simulator` contexts of one mechanism set simulated by one
    persistent launch through a device work queue (``ngpu_batch_*``)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        self.h = C.c_void_p()
        self.members = []
        rc = self.L.ngpu_batch_create(device, C.byref(self.h))
        if rc != 0:
            raise NgpuError(rc, (self.L.ngpu_batch_last_error(None) or b"").decode())

    def _check(self, rc: int) -> int:
        if rc < 0:
            raise NgpuError(rc, (self.L.ngpu_batch_last_error(self.h) or b"").decode())
        return rc

    def add(self, sim: Simulator) -> int:
        idx = self._check(self.L.ngpu_batch_add(self.h, sim._h))
        self.members.append(sim)
        return idx

    def set_launch(self, member: int, waves_per_run: int = 0, block_sync: int = -1, lds_staging: int = -1,
                   runs_per_group: int = 0) -> None:
        cfg = ngpu_batch_launch(waves_per_run, block_sync, lds_staging, runs_per_group)
        self._check(self.L.ngpu_batch_set_launch(self.h, member, C.byref(cfg)))

    def run(self) -> dict:
        st = ngpu_batch_stats()
        self._check(self.L.ngpu_batch_run(self.h, C.byref(st)))
        return st.as_dict()

    def member_info(self, member: int) -> dict:
        info = ngpu_batch_member()
        self._check(self.L.ngpu_batch_member_info(self.h, member, C.byref(info)))
        return info.as_dict()

    def close(self) -> None:
        if self.h and self.h.value:
            self.L.ngpu_batch_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def workload_simulator(w, device: int = 0, mechset: Optional[int] = None) -> Simulator:
    """A configured :class:`Simulator` for a :class:`models.Workload` (``mechset``
    as in :func:`simulate_workload`)."""
    m = w.model
    sim = Simulator(device)
    sim.set_tree(m.tree, m.mechset_id if mechset is None else int(mechset))
    sim.set_sim(m.v_init, m.rec_sites)
    sim.set_params(w.params, m.ncomps, m.nparams)
    sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
    return sim
