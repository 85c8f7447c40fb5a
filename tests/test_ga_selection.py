"""The GA driver's multi-objective parts on the CPU: IBEA (Zitzler & Kuenzli
2004, the default selector of bluepyopt's DEAPOptimisation that
Figures/Figure4_DEAP/optimize_parameters_genetic_alg_par.py:50 runs), the
Pareto-front hall of fame (deap.tools.ParetoFront there), DEAP's varAnd, and
checkpoint/resume (``opt.run(cp_filename=..., cp_frequency=...,
continue_cp=...)``, :52-55).  bluepyopt and deap are not installed, so no
vector from them exists: the checks are known answers worked by hand, the
indicator's dominance properties, a brute-force restatement of the removal
loop, and convergence on ZDT1, whose Pareto front is known in closed form."""
import dataclasses
import itertools

import numpy as np
import pytest

from neurogpu_amd import ga, selection


def test_eps_indicator_known_answers():
    F = np.array([[0.0, 1.0], [1.0, 0.0], [2.0, 2.0]])         # ranges (2, 2)
    I = selection.eps_indicator(F)
    # I[a, b] = max_i (F[a,i] - F[b,i]) / 2
    want = np.array([[0.0, 0.5, -0.5],
                     [0.5, 0.0, -0.5],
                     [1.0, 1.0, 0.0]])
    np.testing.assert_array_equal(I, want)
    # a constant objective has range 0 and counts with range 1:
    # I[0, 1] = max(1 - 2, 3 - 3) = 0, I[1, 0] = max(2 - 1, 0) = 1
    I2 = selection.eps_indicator(np.array([[1.0, 3.0], [2.0, 3.0]]))
    np.testing.assert_array_equal(I2, [[0.0, 0.0], [1.0, 0.0]])


def test_ibea_fitness_known_answer():
    """Two points, a dominating b: I(a,b) = -1, I(b,a) = 1, c = 1, so
    F(a) = -exp(-1/0.05) and F(b) = -exp(1/0.05)."""
    F = np.array([[0.0, 0.0], [1.0, 1.0]])
    fit, T = selection.ibea_fitness(selection.eps_indicator(F))
    np.testing.assert_allclose(fit, [-np.exp(-20.0), -np.exp(20.0)], rtol=1e-15)
    assert T[0, 0] == 0.0 and T[1, 1] == 0.0


def _dominates(a, b):
    return np.all(a <= b) and np.any(a < b)


def test_ibea_fitness_is_dominance_compliant():
    """If a dominates b, F(a) > F(b): every other y has I(y, a) >= I(y, b),
    and a's term in b's sum is below b's term in a's."""
    rng = np.random.default_rng(5)
    for trial in range(30):
        F = rng.integers(0, 6, size=(40, 3)).astype(float)      # many ties and dominations
        fit, _ = selection.ibea_fitness(selection.eps_indicator(F))
        n = 0
        for a, b in itertools.permutations(range(F.shape[0]), 2):
            if _dominates(F[a], F[b]):
                assert fit[a] > fit[b]
                n += 1
        assert n > 0


def _brute_environmental(F, alpha, kappa=0.05):
    """Zitzler & Kuenzli's environmental selection written as its definition:
    repeatedly remove the individual of least fitness, the fitness of every
    survivor recomputed from scratch over the survivors (c and the
    objective ranges fixed by the initial set, as in the paper)."""
    n, m = F.shape
    rng_ = F.max(axis=0) - F.min(axis=0)
    rng_[rng_ == 0] = 1.0
    I = [[max((F[a, i] - F[b, i]) / rng_[i] for i in range(m)) for b in range(n)] for a in range(n)]
    c = max(abs(I[a][b]) for a in range(n) for b in range(n)) or 1.0
    alive = list(range(n))

    def fitness(x):
        return sum(-np.exp(-I[y][x] / (c * kappa)) for y in alive if y != x)
    while len(alive) > alpha:
        f = [fitness(x) for x in alive]
        alive.pop(int(np.argmin(f)))
    return np.array(alive), np.array([fitness(x) for x in alive])


def test_environmental_selection_equals_its_definition():
    rng = np.random.default_rng(6)
    for n, m, alpha in ((12, 2, 5), (20, 3, 11), (9, 4, 1), (16, 2, 16)):
        F = rng.normal(size=(n, m))
        keep, fit = selection.environmental_selection(F, alpha)
        bk, bf = _brute_environmental(F, alpha)
        np.testing.assert_array_equal(keep, bk)
        np.testing.assert_allclose(fit, bf, rtol=1e-12, atol=1e-15)     # incremental sums: rounding residue


def test_environmental_selection_keeps_dominators_of_survivors():
    """A survivor's dominators survive too (their fitness is always higher)."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        F = rng.integers(0, 8, size=(30, 2)).astype(float)
        keep, _ = selection.environmental_selection(F, 12)
        kept = set(keep.tolist())
        for b in kept:
            for a in range(F.shape[0]):
                if _dominates(F[a], F[b]):
                    assert a in kept


def test_tournament_picks_the_fittest_of_each_draw():
    fit = np.array([3.0, -1.0, 7.0, 7.0, 0.5])
    rng = np.random.default_rng(8)
    win = selection.tournament(np.random.default_rng(8), fit, 500, 4)
    cand = rng.integers(fit.size, size=(500, 4))
    for k in range(500):
        c = cand[k]
        assert fit[win[k]] == fit[c].max()
        assert win[k] == c[np.argmax(fit[c])]                     # ties: the first drawn
    # the best of n wins a 4-way tournament with probability 1 - (1 - 1/n)^4
    fit = np.arange(10.0)
    w = selection.tournament(np.random.default_rng(9), fit, 200000, 4)
    assert abs(np.mean(w == 9) - (1 - 0.9 ** 4)) < 0.005


def test_pareto_front_is_the_nondominated_set_of_everything_seen():
    rng = np.random.default_rng(10)
    hof = selection.ParetoFront()
    allx, allf = [], []
    for _ in range(15):
        f = rng.integers(0, 12, size=(25, 3)).astype(float)
        x = rng.normal(size=(25, 4))
        hof.update(x, f)
        allx.append(x)
        allf.append(f)
    X, Fo = np.vstack(allx), np.vstack(allf)
    nd = [i for i in range(Fo.shape[0]) if not any(_dominates(Fo[j], Fo[i]) for j in range(Fo.shape[0]))]
    want = {tuple(r) for r in Fo[nd]}
    got = [tuple(r) for r in hof.f]
    assert len(got) == len(set(got)) == len(want) == len(hof)          # each vector once
    assert set(got) == want
    for x, f in zip(hof.x, hof.f):                                      # rows stay paired
        k = next(i for i in range(Fo.shape[0]) if np.array_equal(X[i], x))
        np.testing.assert_array_equal(Fo[k], f)


def test_var_and_rows():
    rng = np.random.default_rng(11)
    lo, hi = np.full(6, -50.0), np.full(6, 50.0)
    pool = rng.uniform(-1, 1, size=(41, 6))
    # crossover only: each pair keeps its per-gene sum (SBX children are
    # symmetric about the parents' midpoint; far bounds, no clipping); the
    # odd last row is copied
    kids = ga.var_and_rows(np.random.default_rng(1), pool, lo, hi, cxpb=1.0, mutpb=0.0)
    np.testing.assert_allclose(kids[0:40:2] + kids[1:40:2], pool[0:40:2] + pool[1:40:2], atol=1e-12)
    assert not np.array_equal(kids[:40], pool[:40])
    np.testing.assert_array_equal(kids[40], pool[40])
    # neither: copies
    np.testing.assert_array_equal(ga.var_and_rows(rng, pool, lo, hi, cxpb=0.0, mutpb=0.0), pool)
    # the pair's children are sbx_pair_rows' children
    u = np.random.default_rng(1).random((4, 20, 6))            # var_and_rows draws these first
    c1, c2 = ga.sbx_pair_rows(pool[0:40:2], pool[1:40:2], lo, hi, u[0], u[1], u[2])
    np.testing.assert_array_equal(kids[0:40:2], c1)
    np.testing.assert_array_equal(kids[1:40:2], c2)
    # mutation only, every gene: all stay in bounds and move
    lo2, hi2 = np.full(6, -1.0), np.full(6, 1.0)
    k2 = ga.var_and_rows(rng, pool, lo2, hi2, cxpb=0.0, mutpb=1.0, indpb=1.0)
    assert np.all(k2 >= lo2) and np.all(k2 <= hi2) and np.all(k2 != pool)


def test_sbx_pair_rows_second_child_is_the_mirror():
    rng = np.random.default_rng(12)
    lo, hi = np.full(5, 0.0), np.full(5, 1.0)
    A, B = rng.random((50, 5)), rng.random((50, 5))
    u = rng.random((3, 50, 5))
    c1, c2 = ga.sbx_pair_rows(A, B, lo, hi, u[0], u[1], u[2])
    c2_by_swap, _ = ga.sbx_pair_rows(A, B, lo, hi, u[0], u[1], 1.0 - u[2] + (u[2] == 0.5) * 1e-9)
    crossed = u[0] <= 0.5
    np.testing.assert_array_equal(c2[crossed], c2_by_swap[crossed])
    np.testing.assert_array_equal(c2[~crossed], B[~crossed])


# ----------------------------------------------------------------- driver
@dataclasses.dataclass
class _Free:
    lo: float
    hi: float


class _ZDT1:
    """ZDT1 (Zitzler, Deb & Thiele 2000) on genes g in [0, 1]: f1 = g0,
    f2 = h (1 - sqrt(g0 / h)), h = 1 + 9 mean(g1..).  The GA searches log10
    of its values, so values 10**g over [1, 10] give g back.  The Pareto
    front is h = 1: f2 = 1 - sqrt(f1)."""

    def __init__(self, ngenes=6):
        self.m = type("M", (), {"free": [_Free(1.0, 10.0)] * ngenes})()
        self.nevals = 0

    def evaluate(self, x):
        g = np.log10(x)
        h = 1.0 + 9.0 * g[:, 1:].mean(axis=1)
        f1 = g[:, 0]
        self.nevals += x.shape[0]
        return np.stack([f1, h * (1.0 - np.sqrt(f1 / h))], axis=1), g


def _hypervolume2(F, ref=(1.1, 11.0)):
    """Area dominated by the 2-objective set F inside the reference box."""
    P = F[(F[:, 0] < ref[0]) & (F[:, 1] < ref[1])]
    P = P[np.argsort(P[:, 0])]
    area, ylim = 0.0, ref[1]
    for x, y in P:
        if y < ylim:
            area += (ref[0] - x) * (ylim - y)
            ylim = y
    return area


def test_hypervolume_helper():
    assert _hypervolume2(np.array([[0.1, 1.0]]), (1.0, 2.0)) == pytest.approx(0.9)
    assert _hypervolume2(np.array([[0.0, 1.0], [0.5, 0.0]]), (1.0, 2.0)) == pytest.approx(1.0 + 0.5)


def test_ibea_converges_on_zdt1():
    """IBEA's front approaches ZDT1's (h -> 1) and covers more of the
    objective space than the same number of random evaluations."""
    ev = _ZDT1()
    hof = selection.ParetoFront()
    res = ga.run_ga(ev, mu=60, lam=60, ngen=60, seed=3, selector="ibea", hof=hof)
    assert ev.nevals == 60 * 61
    h = 1.0 + 9.0 * res.population[:, 1:].mean(axis=1)
    g = hof.x
    hh = 1.0 + 9.0 * g[:, 1:].mean(axis=1)
    assert np.median(hh) < 1.25, np.median(hh)
    rnd = _ZDT1()
    Fr, _ = rnd.evaluate(10 ** np.random.default_rng(3).uniform(0, 1, size=(ev.nevals, 6)))
    assert _hypervolume2(hof.f) > _hypervolume2(Fr) + 1.0
    assert np.median(h) < 2.0
    # the hall of fame holds no dominated pair
    assert not selection.dominated(hof.f).any()


def _run(ev_factory, tmp, ngen, selector, continue_from=None, cp_every=1, state=None):
    ev = ev_factory()
    hof = selection.ParetoFront()
    kw = dict(mu=24, lam=24, seed=77, selector=selector, hof=hof, cp_filename=str(tmp), cp_frequency=cp_every)
    if state is not None:
        kw["save_state"], kw["load_state"] = state
    res = ga.run_ga(ev, ngen=ngen, continue_cp=continue_from is not None, **kw)
    return res, ev


@pytest.mark.parametrize("selector", ["ibea", "sum"])
def test_resume_from_checkpoint_equals_the_uninterrupted_run(tmp_path, selector):
    whole, ev1 = _run(_ZDT1, tmp_path / "whole.npz", 9, selector)
    cp = tmp_path / "part.npz"
    store = {"w": np.arange(5.0)}
    saved = {}

    def save():
        return {"w": store["w"] * 2}

    def load(d):
        saved.update(d)
    _run(_ZDT1, cp, 4, selector, state=(save, load))
    resumed, ev2 = _run(_ZDT1, cp, 9, selector, continue_from=cp, state=(save, load))
    np.testing.assert_array_equal(saved["w"], np.arange(5.0) * 2)          # caller state came back
    np.testing.assert_array_equal(resumed.population, whole.population)
    np.testing.assert_array_equal(resumed.objectives, whole.objectives)
    np.testing.assert_array_equal(resumed.best_x, whole.best_x)
    assert resumed.best_score == whole.best_score
    np.testing.assert_array_equal(resumed.hof.f, whole.hof.f)
    np.testing.assert_array_equal(resumed.hof.x, whole.hof.x)
    assert ev2.nevals == ev1.nevals
    assert [h[0] for h in resumed.history] == list(range(10))
    assert [h[2] for h in resumed.history] == [h[2] for h in whole.history]


def test_checkpoint_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        ga.run_ga(_ZDT1(), mu=8, lam=8, ngen=2, cp_filename=str(tmp_path / "none.npz"), continue_cp=True)
    cp = tmp_path / "cp.npz"
    ga.run_ga(_ZDT1(), mu=8, lam=8, ngen=2, selector="sum", cp_filename=str(cp))
    with pytest.raises(ValueError):
        ga.run_ga(_ZDT1(), mu=8, lam=8, ngen=4, selector="ibea", cp_filename=str(cp), continue_cp=True)
    with pytest.raises(ValueError):
        ga.run_ga(_ZDT1(), mu=8, lam=16, ngen=2, selector="ibea")
    with pytest.raises(ValueError):
        ga.run_ga(_ZDT1(), mu=8, lam=8, ngen=2, selector="nsga2")
    # the file is plain arrays: it loads without pickle
    with np.load(cp, allow_pickle=False) as z:
        assert int(z["gen"]) == 2 and z["pop"].shape == (8, 6)


def test_environmental_selection_size_limit():
    with pytest.raises(ValueError):
        selection.environmental_selection(np.zeros((selection.IBEA_MAX + 1, 2)), 10)


def test_deap_optimisation_facade(tmp_path):
    """The reference's call shape: DEAPOptimisation(evaluator, seed=1178,
    offspring_size=..., eta=20, mutpb=0.3, cxpb=0.7, hof=ParetoFront()) and
    opt.run(max_ngen=..., cp_filename=..., cp_frequency=1[, continue_cp=True])."""
    cp = str(tmp_path / "cp.npz")
    opt = ga.DEAPOptimisation(_ZDT1(), seed=1178, offspring_size=20, eta=20, mutpb=0.3, cxpb=0.7,
                              hof=selection.ParetoFront())
    pop, hof, log, hist = opt.run(max_ngen=3, cp_filename=cp, cp_frequency=1)
    assert pop.shape == (20, 6) and np.all(pop >= 1.0) and np.all(pop <= 10.0)
    assert [r["gen"] for r in log] == [0, 1, 2, 3] and log[-1]["nevals"] == 80
    assert all(r["min"] <= r["avg"] <= r["max"] for r in log)
    assert len(hof) > 0 and len(hist) == 4
    opt2 = ga.DEAPOptimisation(_ZDT1(), seed=1178, offspring_size=20, eta=20, mutpb=0.3, cxpb=0.7,
                               hof=selection.ParetoFront())
    pop2, _, log2, hist2 = opt2.run(max_ngen=5, cp_filename=cp, cp_frequency=1, continue_cp=True)
    assert [r["gen"] for r in log2] == [4, 5] and len(hist2) == 6
    with pytest.raises(ValueError):
        ga.DEAPOptimisation(_ZDT1(), selector_name="NSGA2")


@pytest.mark.parametrize("selector", ["sum", "ibea"])
def test_init_population_of_simulated_individuals(selector):
    """The screened GA starts from individuals already simulated (genomes,
    their summed scores and objective rows); the best of them is kept."""
    ev = _ZDT1()
    g = np.random.default_rng(4).uniform(0, 1, size=(50, 6))
    obj, _ = ev.evaluate(10 ** g)
    ev.nevals = 0
    res = ga.run_ga(ev, mu=20, lam=20, ngen=3, seed=1, selector=selector, init=(g, obj.sum(axis=1), obj))
    assert ev.nevals == 60
    assert res.best_score <= obj.sum(axis=1).min()
    with pytest.raises(ValueError):
        ga.run_ga(_ZDT1(), mu=20, lam=20, ngen=1, init=(g, obj.sum(axis=1)))
