"""GA layer: variation operators (CPU) and the evaluator / surrogate screen on the GPU."""
import numpy as np
import pytest

from neurogpu_amd import ga


def test_sbx_and_mutation_stay_in_bounds():
    rng = np.random.default_rng(0)
    lo, hi = np.full(8, -3.0), np.full(8, 1.0)
    for _ in range(200):
        a, b = rng.uniform(lo, hi), rng.uniform(lo, hi)
        c1, c2 = ga.sbx(rng, a, b, lo, hi)
        m = ga.poly_mutation(rng, a, lo, hi, indpb=1.0)
        for y in (c1, c2, m):
            assert np.all(y >= lo) and np.all(y <= hi)


def test_sbx_preserves_the_mean_of_each_gene():
    """SBX children are symmetric about the parents' midpoint (Deb & Agrawal 1995)."""
    rng = np.random.default_rng(1)
    lo, hi = np.full(4, -100.0), np.full(4, 100.0)
    a, b = np.array([0.0, 1.0, 2.0, 3.0]), np.array([1.0, -1.0, 5.0, 3.0])
    for _ in range(100):
        c1, c2 = ga.sbx(rng, a, b, lo, hi)
        np.testing.assert_allclose(c1 + c2, a + b, atol=1e-9)


def _scalar_sbx(a, b, lo, hi, ug, ub, us, eta=20.0):
    """ga.sbx's per-gene arithmetic with given uniform numbers, first child."""
    c1 = a.copy()
    for i in range(a.size):
        if ug[i] > 0.5 or abs(a[i] - b[i]) < 1e-14:
            continue
        x1, x2 = min(a[i], b[i]), max(a[i], b[i])
        r = ub[i]
        beta = 1.0 + 2.0 * (x1 - lo[i]) / (x2 - x1)
        alpha = 2.0 - beta ** -(eta + 1)
        bq = (r * alpha) ** (1 / (eta + 1)) if r <= 1 / alpha else (1 / (2 - r * alpha)) ** (1 / (eta + 1))
        y1 = 0.5 * (x1 + x2 - bq * (x2 - x1))
        beta = 1.0 + 2.0 * (hi[i] - x2) / (x2 - x1)
        alpha = 2.0 - beta ** -(eta + 1)
        bq = (r * alpha) ** (1 / (eta + 1)) if r <= 1 / alpha else (1 / (2 - r * alpha)) ** (1 / (eta + 1))
        y2 = 0.5 * (x1 + x2 + bq * (x2 - x1))
        y1, y2 = min(max(y1, lo[i]), hi[i]), min(max(y2, lo[i]), hi[i])
        c1[i] = y2 if us[i] <= 0.5 else y1
    return c1


def test_vectorised_variation_equals_the_per_gene_operators():
    """sbx_rows / mutate_rows compute what sbx / poly_mutation compute, gene by
    gene, from the same uniform numbers."""
    rng = np.random.default_rng(2)
    lo, hi = np.full(12, -2.0), np.full(12, 1.5)
    A, B = rng.uniform(lo, hi, (300, 12)), rng.uniform(lo, hi, (300, 12))
    B[::7, 3] = A[::7, 3]                                   # equal genes are not crossed
    u = rng.random((5, 300, 12))
    cx = ga.sbx_rows(A, B, lo, hi, u[0], u[1], u[2])
    mu = ga.mutate_rows(A, lo, hi, u[3], u[4], indpb=0.5)
    for k in range(300):
        np.testing.assert_allclose(cx[k], _scalar_sbx(A[k], B[k], lo, hi, u[0, k], u[1, k], u[2, k]), rtol=1e-12)
        y = A[k].copy()
        for i in range(12):
            if u[3, k, i] > 0.5:
                continue
            d1, d2 = (y[i] - lo[i]) / (hi[i] - lo[i]), (hi[i] - y[i]) / (hi[i] - lo[i])
            r = u[4, k, i]
            if r < 0.5:
                dq = (2.0 * r + (1.0 - 2.0 * r) * (1.0 - d1) ** 21) ** (1 / 21) - 1.0
            else:
                dq = 1.0 - (2.0 * (1.0 - r) + 2.0 * (r - 0.5) * (1.0 - d2) ** 21) ** (1 / 21)
            y[i] = min(max(y[i] + dq * (hi[i] - lo[i]), lo[i]), hi[i])
        np.testing.assert_allclose(mu[k], y, rtol=1e-12)
    kids = ga.var_or_rows(rng, A, 5000, lo, hi)
    assert kids.shape == (5000, 12) and np.all(kids >= lo) and np.all(kids <= hi)


def test_fill_and_objectives():
    f = np.array([[[2.0, 5.0, 90.0, 1.0, 20.0], [0.0, np.nan, np.nan, np.nan, np.nan]]])
    filled = ga.fill_features(f, 100.0)
    assert filled[0, 1, 1] == 100.0 and filled[0, 1, 2] == 0.0 and filled[0, 1, 4] == 100.0
    tgt = ga.Target(filled[0], np.ones(5))
    assert np.all(ga.objectives(filled, tgt) == 0)


@pytest.mark.gpu
def test_ga_improves_and_surrogate_screen_beats_random():
    """A short GA improves the best simulated score; a surrogate trained on
    simulator features picks offspring whose simulated scores are better than a
    random pick of the same size (the gain the network brings to the GA)."""
    from neurogpu_amd import models
    from neurogpu_amd.surrogate import MLPEnsemble, feature_targets, genome_to_input, make_screen
    w = models.workload("active500", nsets=1, nt=2000)           # 40 ms protocol
    m = w.model
    a, d = models.step_stimulus([0.2, 0.5, 1.0], 2000, 0.02, 5.0, 35.0)
    w.amps, w.dts = a, d
    ev = ga.Evaluator(w, 5.0, 35.0, target_x=m.nominal())
    res = ga.run_ga(ev, mu=64, lam=128, ngen=4, seed=1178)
    assert res.history[-1][2] < res.history[0][2]          # selection found a better simulated score
    lo = np.log10([fp.lo for fp in m.free])
    hi = np.log10([fp.hi for fp in m.free])
    rng = np.random.default_rng(0)
    g = rng.uniform(lo, hi, size=(2048, lo.size))
    f = ev.features(10 ** g)
    ens = MLPEnsemble(lo.size, 3 * len(ga.GA_FEATURES), members=3, seed=1)
    ens.fit(genome_to_input(g, lo, hi), feature_targets(f), epochs=60, batch=256, lr=2e-3)
    cand = rng.uniform(lo, hi, size=(1000, lo.size))
    true = ga.objectives(ev.features(10 ** cand), ev.target).sum(axis=1)
    pick = make_screen(ens, ev.target, lo, hi, 3, keep_frac=0.1, disagree_frac=0.0)(cand)
    rand = rng.choice(cand.shape[0], pick.size, replace=False)
    assert np.median(true[pick]) < np.median(true[rand])
    ens.close()
    ev.close()


@pytest.mark.gpu
def test_ibea_on_the_device_evaluator_resumes_bitwise(tmp_path):
    """bluepyopt's loop (IBEA selector, Pareto-front hall of fame,
    checkpoint every generation) on simulated scores: a run stopped after two
    generations and resumed from its checkpoint ends where the uninterrupted
    run ends (the simulation is deterministic)."""
    from neurogpu_amd import models, selection
    w = models.workload("active500", nsets=1, nt=1500)
    m = w.model
    a, d = models.step_stimulus([0.3, 1.0], 1500, 0.02, 5.0, 25.0)
    w.amps, w.dts = a, d
    ev = ga.Evaluator(w, 5.0, 25.0, target_x=m.nominal())
    print("case IBEA GA on active500, 32 individuals x 4 generations", flush=True)
    runs = []
    for stop, resume in ((4, False), (2, False), (4, True)):
        ev.nevals = 0
        hof = selection.ParetoFront()
        cp = str(tmp_path / ("whole.npz" if stop == 4 and not resume else "part.npz"))
        runs.append(ga.run_ga(ev, mu=32, lam=32, ngen=stop, seed=5, selector="ibea", hof=hof,
                              cp_filename=cp, continue_cp=resume))
    whole, _, resumed = runs
    np.testing.assert_array_equal(resumed.population, whole.population)
    np.testing.assert_array_equal(resumed.objectives, whole.objectives)
    np.testing.assert_array_equal(resumed.hof.f, whole.hof.f)
    assert resumed.best_score == whole.best_score <= whole.history[0][2]
    assert not selection.dominated(whole.hof.f).any() and len(whole.hof) > 0
    ev.close()
