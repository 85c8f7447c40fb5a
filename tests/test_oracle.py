"""Pins for the FP64 CPU oracle (oracle/cpu_ref.c): dense-solve equivalence,
closed forms, invariants, and the shared transcendental functions."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import TREE_FILES, TTPC_DIR, spike_counts
from neurogpu_amd import formats, models, tree
from oracle import cpu_ref


def _matrix(t, u):
    """Tree matrix in Hines-row order: A[h][h] = u[h]; child c of parent p:
    A[c][p] = f[c], A[p][c] = e[c]."""
    A = np.diag(np.asarray(u, dtype=float))
    for h in range(t.N - 1):
        p = t.ks[h + 1] - 1
        A[h, p] = t.f[h]
        A[p, h] = t.e[h]
    return A


@pytest.mark.parametrize("name", ["ttpc512", "mainen384", "cc224", "passive384"])
def test_solve_has_tiny_backward_error(name):
    """The reference trees mix zero-area and large nodes, so the matrices are
    badly scaled and a dense LU is no sharper a reference than the tree solve.
    Pin every mode by its componentwise backward error instead:
    |A x - b|_i <= 1e-13 (sum_j |A_ij x_j| + |b_i|)."""
    t = formats.read_tree(TREE_FILES[name])
    rng = np.random.default_rng(3)
    A0 = _matrix(t, np.zeros(t.N))
    u = np.abs(A0).sum(axis=0) + np.abs(A0).sum(axis=1) + rng.uniform(0.01, 1.0, t.N)
    A = _matrix(t, u)
    b = rng.normal(size=t.N)
    for mode in (0, 1, 2):
        x = cpu_ref.solve(t, u, b, mode)
        r = np.abs(A @ x - b)
        scale = np.abs(A) @ np.abs(x) + np.abs(b)
        assert np.max(r / scale) < 1e-13, mode


@pytest.mark.parametrize("idx,name", list(enumerate(models.WORKLOADS)))
def test_solve_matches_dense_solve_on_generated_trees(idx, name):
    t = models.workload(name, nsets=1, nt=1).model.tree
    rng = np.random.default_rng(idx)
    u = rng.uniform(1.0, 2.0, t.N)
    A0 = _matrix(t, np.zeros(t.N))
    u = u + np.abs(A0).sum(axis=1)
    A = _matrix(t, u)
    b = rng.normal(size=t.N)
    x_ref = np.linalg.solve(A, b)
    for mode in (0, 1, 2):
        np.testing.assert_allclose(cpu_ref.solve(t, u, b, mode), x_ref, rtol=1e-11, atol=1e-12)


def test_beforelu_equals_serial_elimination_bitwise():
    """The level schedule only reorders independent eliminations; with the
    father folds in ascending row order it is bit-identical to eliminating
    rows 0..N-2 into their parents in order."""
    t = formats.read_tree(TREE_FILES["ttpc512"])
    rng = np.random.default_rng(5)
    u = rng.uniform(1.0, 3.0, t.N) + 2.0
    b = rng.normal(size=t.N)
    np.testing.assert_array_equal(cpu_ref.solve(t, u, b, 0), cpu_ref.solve(t, u, b, 1))


def test_single_compartment_backward_euler_closed_form():
    """N = 1 passive compartment: V_{n+1} = V_n + (I - g (V_n - E)) / (C/dt + g)."""
    par = np.array([-1])
    t = tree.build_tree(par, [0.0], [0.0], [1.0], [0], np.ones((1, 1)), 1)
    g, E, dt, nt = 1e-3, -70.0, 0.025, 200
    area = 1000.0
    amps = np.zeros((1, nt))
    amps[0, 20:] = 0.1
    vh = cpu_ref.simulate(t, cpu_ref.MS_PASSIVE, np.array([[g, E]]), amps, np.full(nt, dt), 0, area, -65.0, [0])
    V = -65.0
    for n in range(nt):
        assert abs(vh[0, 0, 0, n] - V) <= 1e-12 * abs(V)
        I = 100 * amps[0, n] / area
        V = V + (I - g * (V - E)) / (1.0 / (dt * 1000) + g)


def test_rest_is_a_fixed_point_bitwise():
    w = models.workload("passive64", nsets=3, nt=400)
    m = w.model
    P = m.allparams(w.x)
    vh = cpu_ref.simulate(m.tree, m.mechset_id, P, np.zeros_like(w.amps), w.dts, m.stim_seg, m.stim_area,
                          -70.0, m.rec_sites)
    assert np.all(vh == -70.0)


def test_identical_instances_identical_traces():
    w = models.workload("active500", nsets=2, nt=600)
    m = w.model
    x = np.vstack([w.x[0], w.x[0]])
    vh = cpu_ref.simulate(m.tree, m.mechset_id, m.allparams(x), w.amps, w.dts, m.stim_seg, m.stim_area,
                          m.v_init, m.rec_sites)
    np.testing.assert_array_equal(vh[0], vh[1])


def test_passive_linearity():
    """A passive cell at rest responds linearly: trace(2I) - rest = 2 (trace(I) - rest)."""
    w = models.workload("passive64", nsets=2, nt=800)
    m = w.model
    P = m.allparams(w.x)
    a1 = w.amps
    v1 = cpu_ref.simulate(m.tree, 0, P, a1, w.dts, m.stim_seg, m.stim_area, -70.0, m.rec_sites)
    v2 = cpu_ref.simulate(m.tree, 0, P, 2 * a1, w.dts, m.stim_seg, m.stim_area, -70.0, m.rec_sites)
    np.testing.assert_allclose(v2 + 70.0, 2 * (v1 + 70.0), atol=1e-9)


def test_dense_step_equivalence_passive():
    """One step of the whole passive assembly equals a dense backward-Euler
    step of the cable system (NEURON discretisation)."""
    w = models.workload("passive64", nsets=1, nt=2)
    m = w.model
    t = m.tree
    g, E = w.x[0, 0], -70.0
    v0 = -70.0
    dt = w.dts[0]
    amps = np.full((1, 2), 0.5)
    vh, vfin = cpu_ref.simulate(t, 0, m.allparams(w.x), amps, w.dts[:1], m.stim_seg, m.stim_area, v0,
                                m.rec_sites, want_vfinal=True)
    N = t.N
    par = t.parent_segments()
    A = np.zeros((N, N))
    rhs = np.zeros(N)
    for s in range(N):
        A[s, s] += t.cm[s] / (dt * 1000) + g
        rhs[s] += -g * (v0 - E)
    rhs[m.stim_seg] += 100 * 0.5 / m.stim_area
    for s in range(1, N):
        p, h = par[s], N - 1 - s
        A[s, s] -= t.f[h]
        A[s, p] += t.f[h]
        A[p, p] -= t.e[h]
        A[p, s] += t.e[h]
    dv = np.linalg.solve(A, rhs)
    np.testing.assert_allclose(vfin[0], v0 + dv, atol=1e-10)


@pytest.mark.parametrize("ms,mech,points", [
    (1, 0, [-27.0]), (1, 3, [-154.9]), (1, 7, [-38.0, -17.0, -64.4]), (1, 8, [-38.0, -66.0]),
    (1, 9, [-32.0, -60.0]),
])
def test_singularity_bumps_are_continuous(ms, mech, points):
    """At the removable singularities the rate code shifts v by 1e-4 mV; the
    result must sit between the values on either side."""
    for v in points:
        mid = cpu_ref.rates(ms, mech, v)
        lo = cpu_ref.rates(ms, mech, v - 2e-4)
        hi = cpu_ref.rates(ms, mech, v + 2e-4)
        assert np.all(np.isfinite(mid))
        for k in range(4):
            a, b = sorted((lo[k], hi[k]))
            assert a - 1e-9 * abs(a) <= mid[k] <= b + 1e-9 * abs(b)


@pytest.mark.parametrize("ms,mech", [(1, m) for m in (0, 1, 3, 4, 5, 6, 7, 8, 9, 12)] +
                         [(2, 0), (2, 3), (2, 4), (2, 5)])
def test_rates_are_gates(ms, mech):
    P = np.zeros(40)
    if ms == 2:
        kin = models._MAINEN_KIN
        names = models.MECHSETS["mainen"]["params"]
        for i, n in enumerate(names):
            P[i] = kin.get(n, 0.0)
    for v in np.linspace(-100, 50, 61):
        r = cpu_ref.rates(ms, mech, v, P)
        assert 0.0 <= r[0] <= 1.0 and 0.0 <= r[2] <= 1.0
        assert r[1] > 0 or (r[1] == 0 and r[0] == 0)
        if r[3] != 0:
            assert r[3] > 0


# oracle rate functions of the TTPC set by mechanism name (cpu_ref.ref_rates indices)
TTPC_RATES = {"Ca_HVA": 0, "Ca_LVAst": 1, "Ih": 3, "Im": 4, "K_Pst": 5, "K_Tst": 6, "Nap_Et2": 7, "NaTa_t": 8,
              "NaTs2_t": 9, "SK_E2": 11, "SKv3_1": 12}


@pytest.mark.parametrize("name", sorted(TTPC_RATES))
def test_ttpc_rates_equal_the_nmodl_sources(name):
    """The oracle's rate functions (restated from the generated
    AllModels.cu) against the same rates evaluated from the mechanisms'
    NMODL PROCEDURE rates blocks (tests/golden/make_mod_rates.py, on a grid
    of 487 voltages plus every singular point, or of calcium values for
    SK_E2): every steady state and time constant agrees to 1e-10 relative.
    A changed constant, a dropped term or a flipped sign moves these by far
    more."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mod_rates_ttpc.npz"))
    keys = list(g[name + "_keys"])
    xs = g["ca"] if name == "SK_E2" else g["v"]
    ref = g[name]
    slot = {"mInf": 0, "mTau": 1, "hInf": 2, "hTau": 3, "zInf": 0}
    got = np.array([[cpu_ref.rates(1, TTPC_RATES[name], float(x))[slot[k]] for k in keys] for x in xs])
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=0)


# Mainen set: oracle rate-function index and the parameter slots (cpu_ref.c's
# layout of a compartment's parameters) that hold each NMODL parameter
MAINEN_RATES = {"ca": (0, {}), "kca": (2, {"caix": 3, "Ra": 4, "Rb": 5}),
                "km": (3, {"tha": 7, "qa": 8, "Ra": 9, "Rb": 10}),
                "kv": (4, {"tha": 12, "qa": 13, "Ra": 14, "Rb": 15}),
                "na": (5, {"tha": 17, "qa": 18, "Ra": 19, "Rb": 20, "thi1": 21, "thi2": 22, "qi": 23, "thinf": 24,
                           "qinf": 25, "Rg": 26, "Rd": 27})}


@pytest.mark.parametrize("name", sorted(MAINEN_RATES))
def test_mainen_rates_equal_the_nmodl_sources(name):
    """The Mainen set's rate functions against its NMODL rates procedures
    (with their efun/trap0 FUNCTIONs, the PARAMETER defaults and celsius =
    37, tests/golden/make_mod_rates.py), the defaults placed in the
    oracle's parameter slots: agreement to 1e-10 relative on 487 voltages
    plus the half-activation points (calcium values for kca)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mod_rates_mainen.npz"))
    mech, slots = MAINEN_RATES[name]
    pars = dict(zip(g[name + "_pnames"], g[name + "_pvalues"]))
    P = np.zeros(40)
    for k, i in slots.items():
        P[i] = pars[k]
    keys = list(g[name + "_keys"])
    slot = {"minf": 0, "mtau": 1, "hinf": 2, "htau": 3, "ninf": 0, "ntau": 1}
    xs = g["cai"] if name == "kca" else g["v"]
    got = np.array([[cpu_ref.rates(2, mech, float(x), P)[slot[k]] for k in keys] for x in xs])
    np.testing.assert_allclose(got, g[name], rtol=1e-10, atol=0)


# oracle slots of each mechanism's BREAKPOINT inputs: (mechanism set, index,
# {NMODL name: ("st", state slot) | ("P", parameter slot)})
CURRENT_SLOTS = {
    "ttpc:Ca_HVA": (1, 0, {"m": ("st", 1), "h": ("st", 2), "gCa_HVAbar": ("P", 0)}),
    "ttpc:Ca_LVAst": (1, 1, {"m": ("st", 3), "h": ("st", 4), "gCa_LVAstbar": ("P", 1)}),
    "ttpc:Ih": (1, 3, {"m": ("st", 5), "gIhbar": ("P", 6), "ehcn": ("P", 7)}),
    "ttpc:Im": (1, 4, {"m": ("st", 6), "gImbar": ("P", 8)}),
    "ttpc:K_Pst": (1, 5, {"m": ("st", 7), "h": ("st", 8), "gK_Pstbar": ("P", 9)}),
    "ttpc:K_Tst": (1, 6, {"m": ("st", 9), "h": ("st", 10), "gK_Tstbar": ("P", 10)}),
    "ttpc:Nap_Et2": (1, 7, {"m": ("st", 11), "h": ("st", 12), "gNap_Et2bar": ("P", 11)}),
    "ttpc:NaTa_t": (1, 8, {"m": ("st", 13), "h": ("st", 14), "gNaTa_tbar": ("P", 12)}),
    "ttpc:NaTs2_t": (1, 9, {"m": ("st", 15), "h": ("st", 16), "gNaTs2_tbar": ("P", 13)}),
    "ttpc:SK_E2": (1, 11, {"z": ("st", 18), "gSK_E2bar": ("P", 16)}),
    "ttpc:SKv3_1": (1, 12, {"m": ("st", 17), "gSKv3_1bar": ("P", 18)}),
    "mainen:ca": (2, 0, {"m": ("st", 1), "h": ("st", 2), "gbar": ("P", 0)}),
    "mainen:kca": (2, 2, {"n": ("st", 4), "gbar": ("P", 2)}),
    "mainen:km": (2, 3, {"n": ("st", 5), "gbar": ("P", 6)}),
    "mainen:kv": (2, 4, {"n": ("st", 6), "gbar": ("P", 11)}),
    "mainen:na": (2, 5, {"m": ("st", 7), "h": ("st", 8), "gbar": ("P", 16)}),
    "mainen:pasx": (2, 6, {"g": ("P", 28), "e": ("P", 29)}),
}


@pytest.mark.parametrize("name", sorted(CURRENT_SLOTS))
def test_currents_equal_the_nmodl_breakpoints(name):
    """Each mechanism's current in the oracle (its term of tt_cur / mn_cur,
    the mechanism alone) against the mechanism's NMODL BREAKPOINT block
    evaluated at 200 random (v, states, conductances, eca) samples
    (tests/golden/make_mod_rates.py, with the models' reversal potentials and
    the Mainen set's tadj at 37 degC): equal to 1e-12 relative, and the
    calcium channels' current is all calcium current."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mod_currents.npz"))
    ms, mech, slots = CURRENT_SLOTS[name]
    inputs = list(g[name + "_inputs"])
    X, want = g[name + "_x"], g[name + "_i"]
    got = np.empty_like(want)
    for j, row in enumerate(X):
        vals = dict(zip(inputs, row))
        st, P = np.zeros(24), np.zeros(40)
        for k, (kind, i) in slots.items():
            (st if kind == "st" else P)[i] = vals[k]
        cur, ica = cpu_ref.current(ms, mech, vals["v"], st, P, vals.get("eca", 120.0))
        got[j] = cur
        if "eca" in vals:
            assert ica == cur
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15 * np.abs(want).max())


# TTPC mechanisms with states: oracle index, state slots, parameter slots
DERIV_SLOTS = {
    "Ca_HVA": (0, {"m": 1, "h": 2}, {}), "Ca_LVAst": (1, {"m": 3, "h": 4}, {}),
    "CaDynamics_E2": (2, {}, {"gamma": 2, "decay": 3, "depth": 4, "minCai": 5}),
    "Ih": (3, {"m": 5}, {}), "Im": (4, {"m": 6}, {}), "K_Pst": (5, {"m": 7, "h": 8}, {}),
    "K_Tst": (6, {"m": 9, "h": 10}, {}), "Nap_Et2": (7, {"m": 11, "h": 12}, {}),
    "NaTa_t": (8, {"m": 13, "h": 14}, {}), "NaTs2_t": (9, {"m": 15, "h": 16}, {}),
    "SK_E2": (11, {"z": 18}, {"zTau": 17}), "SKv3_1": (12, {"m": 17}, {}),
}


@pytest.mark.parametrize("name", sorted(DERIV_SLOTS))
def test_ttpc_state_updates_equal_the_nmodl_derivatives(name):
    """One step of each TTPC mechanism's states in the oracle (the mechanism
    alone) against the exact solution over dt of its NMODL DERIVATIVE block,
    whose equations are linear in their state (tests/golden/
    make_mod_rates.py): 200 random (v, dt, states, cai, ica, parameters)
    samples agree to 1e-11 relative (the two forms of the exponential step
    round differently by up to about 1e-12)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mod_derivs_ttpc.npz"))
    mech, sslots, pslots = DERIV_SLOTS[name]
    inputs, states = list(g[name + "_inputs"]), list(g[name + "_states"])
    X, Y = g[name + "_x"], g[name + "_y"]
    got = np.empty_like(Y)
    for j, row in enumerate(X):
        vals = dict(zip(inputs, row))
        st, P = np.zeros(24), np.zeros(40)
        for k, i in sslots.items():
            st[i] = vals[k]
        for k, i in pslots.items():
            P[i] = vals[k]
        new, cai = cpu_ref.deriv(1, mech, vals["dt"], vals["v"], st, P, vals["cai"], vals["ica"])
        got[j] = [cai if y == "cai" else new[sslots[y]] for y in states]
    np.testing.assert_allclose(got, Y, rtol=1e-11, atol=0)


MAINEN_DERIV_STATES = {"ca": {"m": 1, "h": 2}, "cad": {"ca": 3}, "kca": {"n": 4}, "km": {"n": 5}, "kv": {"n": 6},
                       "na": {"m": 7, "h": 8}}
MAINEN_INDEX = {"ca": 0, "cad": 1, "kca": 2, "km": 3, "kv": 4, "na": 5}


@pytest.mark.parametrize("name", sorted(MAINEN_DERIV_STATES))
def test_mainen_state_updates_equal_the_nmodl_derivatives(name):
    """As for the TTPC set: each Mainen mechanism's update alone against the
    exact solution over dt of its DERIVATIVE block (trates at v + vshift
    with the model's vshift, cad's clamped drive with the model's shell
    depth, the PARAMETER defaults in the oracle's parameter slots), 200
    random samples, 1e-11 relative."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "mod_derivs_mainen.npz"))
    sslots = MAINEN_DERIV_STATES[name]
    pslots = MAINEN_RATES[name][1] if name in MAINEN_RATES else {}
    pars = dict(zip(g[name + "_pnames"], g[name + "_pvalues"]))
    inputs, states = list(g[name + "_inputs"]), list(g[name + "_states"])
    X, Y = g[name + "_x"], g[name + "_y"]
    got = np.empty_like(Y)
    for j, row in enumerate(X):
        vals = dict(zip(inputs, row))
        st, P = np.zeros(24), np.zeros(40)
        for k, i in sslots.items():
            st[i] = vals[k]
        for k, i in pslots.items():
            P[i] = pars[k]
        new, cai = cpu_ref.deriv(2, MAINEN_INDEX[name], vals["dt"], vals["v"], st, P, vals["cai"], vals["ica"])
        got[j] = [new[sslots[y]] for y in states]
        if name == "cad":
            assert cai == new[3]
    np.testing.assert_allclose(got, Y, rtol=1e-11, atol=0)


def test_cnexp_exact_contraction():
    """cnexp is the exact solution at fixed v: y1 - yinf = (y0 - yinf) exp(-dt/tau)."""
    for y0, yinf, tau, dt in [(0.1, 0.7, 3.0, 0.025), (0.9, 0.05, 0.3, 0.1), (0.5, 0.5, 10.0, 0.02)]:
        y1 = cpu_ref.cnexp(y0, yinf, tau, dt)
        assert abs((y1 - yinf) - (y0 - yinf) * math.exp(-dt / tau)) < 1e-15


def test_nernst_constant():
    """ktf/2 ln(cao/cai) at DEF_cai reproduces the reference's def_eca2 constant
    140.2187 mV at 34 C (Code/OneKernel/src/AllModels.cu:36)."""
    assert abs(cpu_ref.ktf(1) / 2 * math.log(2.0 / 5e-5) - 140.218711995) < 1e-4


def test_ttpc_fixture_runs_and_stays_in_bounds():
    """Code/test_output.py: every VHot value lies in (-1000, 1000)."""
    t = formats.read_tree(TREE_FILES["ttpc512"])
    P = formats.read_allparams(TTPC_DIR + "/AllParams.csv", t.ncomps, 19)
    sim = formats.read_sim(TTPC_DIR + "/Sim.csv")
    st = formats.read_stim(TTPC_DIR)
    vh = cpu_ref.simulate(t, 1, P[:4], st.amps, st.dts, st.stim_seg, st.area, sim.v_init, sim.rec_sites)
    assert np.all(np.abs(vh) < 1000) and np.all(np.isfinite(vh))
    assert np.all(spike_counts(vh) > 0)


def _mathf():
    L = cpu_ref.lib()
    L.ref_math.restype = C.c_double
    L.ref_math.argtypes = [C.c_int, C.c_double, C.c_double]
    return L.ref_math


def test_shared_exp_log_within_one_ulp_of_glibc():
    f = _mathf()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-50, 50, 3000), rng.uniform(-740, 705, 3000)])
    worst = max(abs(f(0, x, 0) - math.exp(x)) / math.ulp(math.exp(x)) for x in xs)
    assert worst <= 1.0
    ys = np.concatenate([10 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2, 3000)])
    worst = max(abs(f(1, y, 0) - math.log(y)) / math.ulp(math.log(y)) for y in ys)
    assert worst <= 1.0
    assert f(1, 0.0, 0) == -math.inf and f(1, math.inf, 0) == math.inf and f(0, 800.0, 0) == math.inf
    assert f(0, -800.0, 0) == 0.0 and math.isnan(f(0, math.nan, 0))


def test_shared_pow_relative_error():
    f = _mathf()
    for x in 10 ** np.linspace(-3, 4, 400):
        assert abs(f(2, x, 4.8) / x ** 4.8 - 1) < 2e-14
    assert f(2, 3.0, 2.0) == 9.0 and f(2, 0.7, 4.0) == (0.7 * 0.7) * (0.7 * 0.7)
