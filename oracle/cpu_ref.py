"""ctypes front end of the FP64 CPU oracle (``oracle/cpu_ref.c``).

TEST INFRASTRUCTURE ONLY: used by ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` as the checker; never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NGPU_CPUREF_LIB") or os.path.join(_HERE, "_build", "libcpuref.so")

MS_PASSIVE, MS_TTPC, MS_MAINEN = 0, 1, 2
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)


class _RefTree(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ncomps", C.c_int), ("mechset", C.c_int), ("nmodels", C.c_int),
        ("e", _DP), ("f", _DP), ("cm", _DP),
        ("ks", _IP), ("seg_to_comp", _IP), ("bool_model", _IP), ("son_no", _IP),
        ("depth", C.c_int), ("logn_depth", C.c_int), ("n_fathers", C.c_int), ("n_call", C.c_int),
        ("rel_starts", _IP), ("rel_ends", _IP), ("rel_vec", _IP),
        ("seg_start", _IP), ("seg_end", _IP), ("fathers", _IP), ("fidxs", _IP),
        ("comp_depth", C.c_int), ("comp_fdepth", C.c_int),
        ("comp_by_level32", _IP), ("comp_by_flevel32", _IP),
        ("n_lrel", C.c_int), ("n_flrel", C.c_int),
        ("lrel_starts", _IP), ("lrel_ends", _IP), ("flrel_starts", _IP), ("flrel_ends", _IP),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ref_simulate.restype = C.c_int
        L.ref_simulate.argtypes = [C.POINTER(_RefTree), _DP, C.c_int, _DP, C.c_int, C.c_int,
                                   _DP, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _IP,
                                   _DP, _DP, C.c_int, C.c_int]
        L.ref_solve.restype = C.c_int
        L.ref_solve.argtypes = [C.POINTER(_RefTree), _DP, _DP, C.c_int]
        L.ref_rates.restype = C.c_int
        L.ref_rates.argtypes = [C.c_int, C.c_int, C.c_double, _DP, _DP]
        L.ref_cnexp.restype = C.c_double
        L.ref_current.argtypes = [C.c_int, C.c_int, C.c_double, _DP, _DP, C.c_double, _DP]
        L.ref_current.restype = C.c_int
        L.ref_deriv.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _DP, _DP, C.POINTER(C.c_double), C.c_double]
        L.ref_deriv.restype = C.c_int
        L.ref_cnexp.argtypes = [C.c_double] * 4
        L.ref_ktf.restype = C.c_double
        L.ref_ktf.argtypes = [C.c_int]
        L.ref_mechset_nparams.restype = C.c_int
        L.ref_mechset_nparams.argtypes = [C.c_int]
        _lib = L
    return _lib


class RefTree:
    """Owns the int32/float64 copies the C struct points into."""

    def __init__(self, t, mechset: int):
        keep = self._keep = []

        def ia(x):
            a = np.ascontiguousarray(np.asarray(x).ravel(), dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(_IP)

        def da(x):
            a = np.ascontiguousarray(np.asarray(x).ravel(), dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(_DP)

        self.t = t
        self.s = _RefTree(
            t.N, t.ncomps, mechset, t.nmodels, da(t.e), da(t.f), da(t.cm),
            ia(t.ks), ia(t.seg_to_comp), ia(t.bool_model), ia(t.son_no),
            t.depth, t.logn_depth, t.n_fathers, t.n_call,
            ia(t.rel_starts), ia(t.rel_ends), ia(t.rel_vec), ia(t.seg_start), ia(t.seg_end),
            ia(t.fathers), ia(t.fidxs), t.comp_by_level32.shape[0], t.comp_by_flevel32.shape[0],
            ia(t.comp_by_level32), ia(t.comp_by_flevel32), t.lrel_starts.size, t.flrel_starts.size,
            ia(t.lrel_starts), ia(t.lrel_ends), ia(t.flrel_starts), ia(t.flrel_ends))


def simulate(t, mechset, params, amps, dts, stim_seg, stim_area, v_init, rec_sites,
             solve_mode=0, nthreads=None, want_vfinal=False):
    """Run the oracle.  Returns vhot [nsets, nstim, nrec, nt] (and v_final
    [nsets*nstim, N] when asked)."""
    L = lib()
    rt = RefTree(t, mechset)
    params = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    amps = np.ascontiguousarray(np.atleast_2d(amps), dtype=np.float64)
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    rec = np.ascontiguousarray(rec_sites, dtype=np.int32)
    nsets, nstim, nt = params.shape[0], amps.shape[0], dts.size
    vhot = np.zeros((nsets, nstim, rec.size, nt), dtype=np.float64)
    vfin = np.zeros((nsets * nstim, t.N), dtype=np.float64) if want_vfinal else None
    if nthreads is None:
        try:
            nthreads = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            nthreads = os.cpu_count() or 1
        env = os.environ.get("OMP_NUM_THREADS", "")
        if env.isdigit() and int(env) > 0:
            nthreads = min(nthreads, int(env))
    rc = L.ref_simulate(C.byref(rt.s), params.ctypes.data_as(_DP), nsets,
                        amps.ctypes.data_as(_DP), nstim, amps.shape[1],
                        dts.ctypes.data_as(_DP), nt, int(stim_seg), float(stim_area),
                        float(v_init), rec.size, rec.ctypes.data_as(_IP),
                        vhot.ctypes.data_as(_DP),
                        vfin.ctypes.data_as(_DP) if vfin is not None else None,
                        int(solve_mode), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"ref_simulate failed with code {rc}")
    return (vhot, vfin) if want_vfinal else vhot


def solve(t, u, b, mode=0, mechset=0):
    """Solve one Hines system (u diagonal, b rhs, Hines-row order); returns x."""
    L = lib()
    rt = RefTree(t, mechset)
    uu = np.zeros(t.N + 1)
    bb = np.zeros(t.N + 1)
    uu[: t.N] = u
    bb[: t.N] = b
    rc = L.ref_solve(C.byref(rt.s), uu.ctypes.data_as(_DP), bb.ctypes.data_as(_DP), int(mode))
    if rc != 0:
        raise RuntimeError("ref_solve failed")
    return bb[: t.N].copy()


def rates(mechset, mech, x, params=None):
    L = lib()
    out = np.zeros(4)
    P = np.ascontiguousarray(params if params is not None else np.zeros(40), dtype=np.float64)
    rc = L.ref_rates(mechset, mech, float(x), P.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
    if rc != 0:
        raise ValueError("no rate function for that mechanism")
    return out


def current(mechset, mech, v, st, params, eca=120.0):
    """(current, calcium part) of one mechanism alone: st by state slot,
    params by parameter slot."""
    L = lib()
    out = np.zeros(2)
    S = np.ascontiguousarray(st, dtype=np.float64)
    P = np.ascontiguousarray(params, dtype=np.float64)
    rc = L.ref_current(mechset, mech, float(v), S.ctypes.data_as(_DP), P.ctypes.data_as(_DP), float(eca),
                       out.ctypes.data_as(_DP))
    if rc != 0:
        raise ValueError("no current term for that mechanism")
    return out


def deriv(mechset, mech, dt, v, st, params, cai=5e-5, ica=0.0):
    """One mechanism's state update alone: (new states by slot, new cai)."""
    L = lib()
    S = np.array(st, dtype=np.float64)
    P = np.ascontiguousarray(params, dtype=np.float64)
    c = C.c_double(cai)
    rc = L.ref_deriv(mechset, mech, float(dt), float(v), S.ctypes.data_as(_DP), P.ctypes.data_as(_DP), C.byref(c),
                     float(ica))
    if rc != 0:
        raise ValueError("no state update for that mechanism")
    return S, c.value


def cnexp(y, yinf, tau, dt):
    return lib().ref_cnexp(y, yinf, tau, dt)


def ktf(mechset):
    return lib().ref_ktf(mechset)
