"""The device mechanism code (neurogpu_amd/csrc/mechs.h) evaluates the
reference's rate equations in a rearranged form: rates 1/tau instead of tau,
shared exponentials, reciprocal constants.  Compiled for the host, every rate
function must agree with the oracle's restatement of the reference's forms
(oracle/cpu_ref.c) to a few ulp-scale relative errors, across the voltage
range, at the singularity bumps and at the branch points."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from neurogpu_amd import models
from oracle import cpu_ref

_DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mechs") / "libmechs_host.so")
    src = os.path.join(ROOT, "tests", "mechs_host.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "neurogpu_amd", "csrc"), src, "-o", out], check=True)
    L = C.CDLL(out)
    L.host_rates.restype = C.c_int
    L.host_rates.argtypes = [C.c_int, C.c_int, C.c_double, _DP, _DP]
    L.host_gate.restype = C.c_double
    L.host_gate.argtypes = [C.c_double] * 4
    L.host_expc.restype = C.c_double
    L.host_expc.argtypes = [C.c_double]
    L.host_logc.restype = C.c_double
    L.host_logc.argtypes = [C.c_double]
    return L


def _params(ms):
    P = np.zeros(40)
    if ms == 2:
        kin = models._MAINEN_KIN
        for i, n in enumerate(models.MECHSETS["mainen"]["params"]):
            P[i] = kin.get(n, 0.0)
    return P


def _host(L, ms, mech, v, P):
    out = np.zeros(4)
    assert L.host_rates(ms, mech, float(v), P.ctypes.data_as(_DP), out.ctypes.data_as(_DP)) == 0
    return out


# a dense sweep and the K_Pst branch point v + 10 = -50, then the
# singularity bumps and points next to them
_SWEEP = np.linspace(-130.0, 70.0, 2001) + 0.00731   # off the singular points (see below)
_SPECIAL = [-60.0, -60.0 - 1e-12, -27.0, -154.9, -38.0, -17.0, -64.4, -66.0, -32.0, -60.0, -60.0 - 1e-9, -38.0 + 1e-7, -32.0 + 3e-6,
            -17.0 - 2e-6, -64.4 + 1e-6, -60.000001, -59.999999]

TTPC = [0, 1, 3, 4, 5, 6, 7, 8, 9, 12]
MAINEN = [0, 3, 4, 5]


def _worst(L, ms, mech, vs):
    P = _params(ms)
    worst_inf = worst_rate = 0.0
    for v in vs:
        ref = cpu_ref.rates(ms, mech, v, P)
        got = _host(L, ms, mech, v, P)
        assert np.all(np.isfinite(got))
        for k in (0, 2):
            worst_inf = max(worst_inf, abs(got[k] - ref[k]))
            if ref[k + 1] != 0.0:
                worst_rate = max(worst_rate, abs(got[k + 1] * ref[k + 1] - 1.0))
            else:
                assert got[k + 1] == 0.0
    return worst_inf, worst_rate


@pytest.mark.parametrize("ms,mech", [(1, m) for m in TTPC] + [(2, m) for m in MAINEN])
def test_rates_match_reference_forms(host, ms, mech):
    """Over the sweep: inf to 1e-13 absolute (gates lie in [0, 1]) and
    rate * tau = 1 to 1e-12: a few ulp, widened where z / (exp(z) - 1) is
    evaluated just above the efun threshold |z| = 1e-4, whose cancellation
    both forms share."""
    wi, wr = _worst(host, ms, mech, _SWEEP)
    assert wi < 1e-13, wi
    assert wr < 1e-12, wr


@pytest.mark.parametrize("ms,mech", [(1, m) for m in TTPC])
def test_rates_next_to_singularities(host, ms, mech):
    """Within 1e-9..1e-6 mV of a removable singularity both forms divide by
    1 - exp(x/k) with x/k down to 1e-10, so each carries a rounding error of
    eps * k / |x|; they must still agree to that level (1e-5 relative at the
    closest point) and stay finite."""
    wi, wr = _worst(host, ms, mech, _SPECIAL)
    assert wi < 1e-5, wi
    assert wr < 1e-5, wr


def test_sk_e2_and_kca_calcium_gates(host):
    P = _params(2)
    for ca in np.concatenate([np.logspace(-8, -2, 400), [1e-7, 0.99e-7, 5e-5]]):
        assert abs(_host(host, 1, 11, ca, P)[0] - cpu_ref.rates(1, 11, ca, P)[0]) < 1e-14
        ref = cpu_ref.rates(2, 2, ca, P)
        got = _host(host, 2, 2, ca, P)
        assert abs(got[0] - ref[0]) < 1e-14 and abs(got[1] * ref[1] - 1.0) < 1e-13


def test_gate_update_equals_cnexp(host):
    """y + (1 - e^{-dt rate}) (yinf - y) against the reference's cnexp form."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        y, yinf = rng.uniform(0, 1, 2)
        tau = 10 ** rng.uniform(-2, 3)
        dt = rng.choice([0.02, 0.025, 0.1])
        assert abs(host.host_gate(y, yinf, 1.0 / tau, dt) - cpu_ref.cnexp(y, yinf, tau, dt)) < 1e-15


def test_clamped_exp(host):
    """ng_expc (Estrin-evaluated polynomial) is within one ulp of ng_exp and of
    libm inside [-708, 708] and saturates to finite normal numbers outside it."""
    L = cpu_ref.lib()
    L.ref_math.restype = C.c_double
    L.ref_math.argtypes = [C.c_int, C.c_double, C.c_double]
    xs = np.concatenate([np.linspace(-700, 700, 3001), np.random.default_rng(3).uniform(-60, 60, 5000),
                         [-708.0, 708.0, 1e-300, -1e-300]])
    for x in xs:
        got = host.host_expc(x)
        assert abs(got - L.ref_math(0, x, 0.0)) <= math.ulp(got) and abs(got - math.exp(x)) <= math.ulp(got)
    assert host.host_expc(0.0) == 1.0
    assert host.host_expc(1e6) == host.host_expc(708.0) and math.isfinite(host.host_expc(1e6))
    assert host.host_expc(-1e6) == host.host_expc(-708.0) > 0.0


def test_branch_free_log(host):
    """ng_logc (no special-case branches, Estrin-evaluated atanh series) is
    within one ulp of libm's log over the normal range and exact at 1."""
    ys = np.concatenate([np.logspace(-300, 300, 4001), np.random.default_rng(4).uniform(0.5, 2.0, 5000),
                         np.logspace(-8, 8, 4001), [2.0 ** -1022, 2.0, 0.5, math.sqrt(0.5), math.sqrt(2.0)]])
    for y in ys:
        ref = math.log(y)
        assert abs(host.host_logc(y) - ref) <= math.ulp(ref), y
    assert host.host_logc(1.0) == 0.0
