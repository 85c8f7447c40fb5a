"""Genetic-algorithm layer on top of the simulator: the evaluator (params ->
simulation -> features -> objectives), a plain (mu + lambda) GA, and the
surrogate-screened GA.

Reference behaviour being replaced:
  * evaluator   Figures/Figure4_DEAP/neurogpu_multistim_evaluator_par.py:239-290
                (AllParams.csv -> one neuroGPU process per stimulus -> eFEL
                features per individual -> objectives, NaN -> large penalty)
  * GA loop     PYS/DEAP_helper.py:126-294 (varOr, eaMuPlusLambda) driven by
                bluepyopt DEAPOptimisation(seed=1178, eta=20, cxpb=0.7, mutpb=0.3)
                (Figures/Figure4_DEAP/optimize_parameters_genetic_alg_par.py:50)
Here one libneurogpu call simulates the whole population x all stimuli and
the features are computed on the device.  Variation uses the bounded
simulated-binary crossover and polynomial mutation (eta = 20) that
bluepyopt's DEAP setup uses; selection keeps the mu best by total objective,
or is IBEA (the reference's default, :mod:`selection`); a Pareto-front hall
of fame and checkpoint/resume follow the reference's driver.
"""
from __future__ import annotations

import dataclasses
import json
import os
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from ._lib import Simulator
from .selection import ParetoFront, eps_indicator, ibea_fitness, tournament

# Objective of an undefined feature difference, and of every objective of a
# diverged run (a trace with a NaN sample).  The reference's evaluator scores
# them 10000 and 100000 on raw feature differences (PYS/efel_ext.py:73-83,
# :90-95); on these scale-normalised objectives the same two levels, ten
# times apart, are 250 (also the cap of a finite difference) and 2500.
PENALTY = 250.0
DIVERGED_PENALTY = 2500.0
GA_FEATURES = (0, 1, 2, 4)   # spike count, first-spike latency, AP amplitude, mean ISI per stimulus


def diverged(f: np.ndarray) -> np.ndarray:
    """Runs whose trace had a NaN sample (raw features: spike count NaN)."""
    return np.isnan(np.asarray(f)[..., 0])


def fill_features(f: np.ndarray, window: float) -> np.ndarray:
    """Features with the undefined entries replaced by their natural limits
    (no spike: latency and ISI = the stimulus window, amplitude = width = 0;
    a diverged run: no spike), so that they can be regressed and compared."""
    out = np.array(f, dtype=np.float64, copy=True)
    out[..., 0][np.isnan(out[..., 0])] = 0.0
    lat, amp, wid, isi = (out[..., 1], out[..., 2], out[..., 3], out[..., 4])
    lat[np.isnan(lat)] = window
    amp[np.isnan(amp)] = 0.0
    wid[np.isnan(wid)] = 0.0
    isi[np.isnan(isi)] = window
    return out


@dataclasses.dataclass
class Target:
    features: np.ndarray      # [nstim, NFEAT] filled target features
    scale: np.ndarray         # [NFEAT] per-feature scale of the objectives


def objectives(filled: np.ndarray, target: Target, bad: Optional[np.ndarray] = None) -> np.ndarray:
    """|f - f*| / scale per stimulus and GA feature -> [n, nstim * len(GA_FEATURES)],
    capped at PENALTY; ``bad`` [n, nstim]: diverged runs, whose objectives
    are all DIVERGED_PENALTY."""
    d = np.abs(filled[..., GA_FEATURES] - target.features[None, :, GA_FEATURES]) / target.scale[list(GA_FEATURES)]
    d = np.minimum(np.where(np.isfinite(d), d, PENALTY), PENALTY)
    if bad is not None:
        d = np.where(np.asarray(bad)[..., None], DIVERGED_PENALTY, d)
    return d.reshape(filled.shape[0], -1)


class Evaluator:
    """Simulates a population on one GPU context and scores it."""

    def __init__(self, workload, t_on: float, t_off: float, device: int = 0, target_x: Optional[np.ndarray] = None):
        self.w = workload
        self.m = workload.model
        self.t_on, self.t_off = t_on, t_off
        self.window = t_off - t_on
        self.sim = Simulator(device)
        self.sim.set_tree(self.m.tree, self.m.mechset_id)
        self.sim.set_sim(self.m.v_init, self.m.rec_sites)
        self.sim.set_stimuli(workload.amps, workload.dts, self.m.stim_seg, self.m.stim_area)
        # the parameter table is assembled on the device from each
        # generation's free parameters (no host-side parameter matrix)
        self.sim.set_param_template(self.m.param_template(), self.m.ncomps, self.m.nparams, len(self.m.free))
        self.nevals = 0
        self.sim_seconds = 0.0
        self.target = None
        if target_x is not None:
            f = self.features(np.atleast_2d(target_x))[0]
            self.target = Target(f, np.array([1.0, 2.0, 5.0, 0.2, 2.0]))

    def features(self, x: np.ndarray) -> np.ndarray:
        """Filled features [n, nstim, NFEAT] of free-parameter rows x."""
        x = np.atleast_2d(x)
        t0 = time.perf_counter()
        self.sim.set_genomes(x)
        self.sim.run(fetch=False)
        f = self.sim.features(self.t_on, self.t_off)
        self.sim_seconds += time.perf_counter() - t0
        self.nevals += x.shape[0]
        self.last_diverged = diverged(f)
        return fill_features(f, self.window)

    def evaluate(self, x: np.ndarray):
        f = self.features(x)
        return objectives(f, self.target, self.last_diverged), f

    def close(self):
        self.sim.close()


# ------------------------------------------------------------------ variation
def sbx(rng, a, b, lo, hi, eta=20.0):
    """Bounded simulated-binary crossover (Deb & Agrawal 1995), per gene with p=0.5."""
    c1, c2 = a.copy(), b.copy()
    for i in range(a.size):
        if rng.random() > 0.5 or abs(a[i] - b[i]) < 1e-14:
            continue
        x1, x2 = min(a[i], b[i]), max(a[i], b[i])
        xl, xu = lo[i], hi[i]
        r = rng.random()
        beta = 1.0 + 2.0 * (x1 - xl) / (x2 - x1)
        alpha = 2.0 - beta ** -(eta + 1)
        bq = (r * alpha) ** (1 / (eta + 1)) if r <= 1 / alpha else (1 / (2 - r * alpha)) ** (1 / (eta + 1))
        y1 = 0.5 * (x1 + x2 - bq * (x2 - x1))
        beta = 1.0 + 2.0 * (xu - x2) / (x2 - x1)
        alpha = 2.0 - beta ** -(eta + 1)
        bq = (r * alpha) ** (1 / (eta + 1)) if r <= 1 / alpha else (1 / (2 - r * alpha)) ** (1 / (eta + 1))
        y2 = 0.5 * (x1 + x2 + bq * (x2 - x1))
        y1, y2 = min(max(y1, xl), xu), min(max(y2, xl), xu)
        if rng.random() <= 0.5:
            c1[i], c2[i] = y2, y1
        else:
            c1[i], c2[i] = y1, y2
    return c1, c2


def poly_mutation(rng, a, lo, hi, eta=20.0, indpb=None):
    """Bounded polynomial mutation (Deb 2001), each gene with probability 1/n."""
    y = a.copy()
    n = a.size
    indpb = 1.0 / n if indpb is None else indpb
    for i in range(n):
        if rng.random() > indpb:
            continue
        xl, xu = lo[i], hi[i]
        d1, d2 = (y[i] - xl) / (xu - xl), (xu - y[i]) / (xu - xl)
        r = rng.random()
        mp = 1.0 / (eta + 1.0)
        if r < 0.5:
            xy = 1.0 - d1
            val = 2.0 * r + (1.0 - 2.0 * r) * xy ** (eta + 1)
            dq = val ** mp - 1.0
        else:
            xy = 1.0 - d2
            val = 2.0 * (1.0 - r) + 2.0 * (r - 0.5) * xy ** (eta + 1)
            dq = 1.0 - val ** mp
        y[i] = min(max(y[i] + dq * (xu - xl), xl), xu)
    return y


def sbx_pair_rows(A, B, lo, hi, u_gene, u_beta, u_swap, eta=20.0):
    """Both children of bounded SBX for rows of parents A, B [n, g], from
    given uniform numbers [n, g] (gene crossed when u_gene <= 0.5, spread
    from u_beta, order from u_swap): the arithmetic of :func:`sbx`, one array
    operation per step instead of one Python step per gene."""
    x1, x2 = np.minimum(A, B), np.maximum(A, B)
    cross = (u_gene <= 0.5) & (np.abs(A - B) >= 1e-14)
    span = np.where(cross, x2 - x1, 1.0)
    e1 = eta + 1.0

    def spread(beta):
        alpha = 2.0 - beta ** -e1
        inner = np.where(u_beta <= 1.0 / alpha, u_beta * alpha, 1.0 / (2.0 - u_beta * alpha))
        return inner ** (1.0 / e1)
    y1 = 0.5 * (x1 + x2 - spread(1.0 + 2.0 * (x1 - lo) / span) * span)
    y2 = 0.5 * (x1 + x2 + spread(1.0 + 2.0 * (hi - x2) / span) * span)
    y1, y2 = np.clip(y1, lo, hi), np.clip(y2, lo, hi)
    swap = u_swap <= 0.5
    return np.where(cross, np.where(swap, y2, y1), A), np.where(cross, np.where(swap, y1, y2), B)


def sbx_rows(A, B, lo, hi, u_gene, u_beta, u_swap, eta=20.0):
    """First children of :func:`sbx_pair_rows`."""
    return sbx_pair_rows(A, B, lo, hi, u_gene, u_beta, u_swap, eta)[0]


def mutate_rows(A, lo, hi, u_pick, u_r, eta=20.0, indpb=None):
    """Bounded polynomial mutation of rows A [n, g] from given uniform numbers
    (gene mutated when u_pick <= indpb, default 1/g): the arithmetic of
    :func:`poly_mutation`, vectorised."""
    indpb = 1.0 / A.shape[1] if indpb is None else indpb
    d1, d2 = (A - lo) / (hi - lo), (hi - A) / (hi - lo)
    mp = 1.0 / (eta + 1.0)
    lower = u_r < 0.5
    val = np.where(lower, 2.0 * u_r + (1.0 - 2.0 * u_r) * (1.0 - d1) ** (eta + 1),
                   2.0 * (1.0 - u_r) + 2.0 * (u_r - 0.5) * (1.0 - d2) ** (eta + 1))
    dq = np.where(lower, val ** mp - 1.0, 1.0 - val ** mp)
    y = np.clip(A + dq * (hi - lo), lo, hi)
    return np.where(u_pick <= indpb, y, A)


def var_or_rows(rng, pop, lam, lo, hi, cxpb=0.7, mutpb=0.3, eta=20.0):
    """DEAP's varOr (PYS/DEAP_helper.py:126-) for lam offspring at once: each
    is a crossover child (probability cxpb), a mutant (mutpb) or a copy."""
    n, g = pop.shape
    r = rng.random(lam)
    i = rng.integers(n, size=lam)
    j = (i + 1 + rng.integers(n - 1, size=lam)) % n          # a second, distinct parent
    u = rng.random((5, lam, g))
    cx = sbx_rows(pop[i], pop[j], lo, hi, u[0], u[1], u[2], eta)
    mu = mutate_rows(pop[i], lo, hi, u[3], u[4], eta)
    return np.where((r < cxpb)[:, None], cx, np.where((r < cxpb + mutpb)[:, None], mu, pop[i]))


def var_and_rows(rng, pool, lo, hi, cxpb=0.7, mutpb=0.3, eta=20.0, indpb=0.5):
    """DEAP's varAnd, the variation of bluepyopt's DEAP loop, for a mating pool
    [n, g]: consecutive rows (0,1), (2,3), ... are crossed with probability
    cxpb (both children replace the pair), then every row is mutated with
    probability mutpb (each gene with probability indpb, bluepyopt's 0.5).
    Returns n offspring."""
    n, g = pool.shape
    kids = pool.copy()
    npair = n // 2
    if npair:
        A, B = kids[0:2 * npair:2], kids[1:2 * npair:2]
        u = rng.random((4, npair, g))
        c1, c2 = sbx_pair_rows(A, B, lo, hi, u[0], u[1], u[2], eta)
        do = (u[3][:, 0] < cxpb)[:, None]
        kids[0:2 * npair:2] = np.where(do, c1, A)
        kids[1:2 * npair:2] = np.where(do, c2, B)
    u = rng.random((3, n, g))
    mut = mutate_rows(kids, lo, hi, u[0], u[1], eta, indpb)
    return np.where((u[2][:, 0] < mutpb)[:, None], mut, kids)


def var_or(rng, pop, lam, lo, hi, cxpb=0.7, mutpb=0.3):
    """Offspring as DEAP's varOr (PYS/DEAP_helper.py:126-): crossover, mutation or reproduction."""
    out = np.empty((lam, pop.shape[1]))
    for k in range(lam):
        r = rng.random()
        if r < cxpb:
            i, j = rng.choice(pop.shape[0], 2, replace=False)
            out[k] = sbx(rng, pop[i], pop[j], lo, hi)[0]
        elif r < cxpb + mutpb:
            out[k] = poly_mutation(rng, pop[rng.integers(pop.shape[0])], lo, hi)
        else:
            out[k] = pop[rng.integers(pop.shape[0])]
    return out


@dataclasses.dataclass
class GAResult:
    best_x: np.ndarray
    best_score: float
    history: List[tuple]      # (generation, wall seconds, best score, evaluations)
    population: Optional[np.ndarray] = None     # final population (log10 genomes)
    objectives: Optional[np.ndarray] = None     # their objective rows
    hof: Optional[ParetoFront] = None


def _save_checkpoint(path, selector, gen, rng, pop, obj, best, hist, hof, nevals, extra):
    """A checkpoint as a .npz of plain arrays (no pickle): the population or
    mating pool and its objectives, the best individual, the hall of fame, the
    history, the generator state (JSON), and the caller's arrays (``extra``,
    e.g. surrogate weights)."""
    arrays = {"selector": np.array(selector), "gen": np.array(gen), "pop": pop, "obj": obj,
              "best_x": best[0], "best_score": np.array(best[1]), "nevals": np.array(nevals),
              "history": np.array(hist, dtype=np.float64).reshape(-1, 4),
              "rng_state": np.array(json.dumps(rng.bit_generator.state))}
    if hof is not None:
        arrays["hof_x"], arrays["hof_f"] = hof.x, hof.f
    for k, v in (extra or {}).items():
        arrays["extra_" + k] = np.asarray(v)
    tmp = path + ".tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, path)


def _load_checkpoint(path):
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["rng_state"] = json.loads(str(d["rng_state"]))
    d["extra"] = {k[6:]: v for k, v in d.items() if k.startswith("extra_")}
    return d


def run_ga(evaluator: Evaluator, mu: int, lam: int, ngen: int, seed: int = 1178,
           screen: Optional[Callable] = None, oversample: int = 1, time_limit: Optional[float] = None,
           stop_score: Optional[float] = None, log: Optional[Callable] = None,
           observe: Optional[Callable] = None, n_init: Optional[int] = None,
           init: Optional[tuple] = None, selector: str = "sum", hof: Optional[ParetoFront] = None,
           cp_filename: Optional[str] = None, cp_frequency: int = 1, continue_cp: bool = False,
           save_state: Optional[Callable[[], Dict[str, np.ndarray]]] = None,
           load_state: Optional[Callable[[Dict[str, np.ndarray]], None]] = None,
           tournament_n: int = 4, cxpb: float = 0.7, mutpb: float = 0.3, eta: float = 20.0,
           record: Optional[Callable[[int, int, np.ndarray], None]] = None) -> GAResult:
    """A GA in log10 parameter space.

    ``selector="sum"``: (mu + lambda) -- lam offspring by varOr, the mu of
    least summed objective survive.  With ``screen`` the GA draws
    ``oversample * lam`` offspring and ``screen(genomes)`` returns the indices
    of those to simulate (the surrogate pre-screen); every score that enters
    selection comes from the simulator.  ``n_init`` random individuals
    (default ``mu``) are simulated first and the best ``mu`` of them form the
    first population; ``init`` = (genomes, simulated scores[, objective
    rows) gives the first population instead (individuals already simulated,
    e.g. the surrogate's training set).

    ``selector="ibea"``: the loop of bluepyopt's DEAPOptimisation with its
    default IBEA selector (Figures/Figure4_DEAP/optimize_parameters_genetic_alg_par.py:50,
    ``offspring_size`` = mu, eta 20, cxpb 0.7, mutpb 0.3): a mating pool of mu
    (first the mu random individuals), offspring by varAnd, and the next pool
    by ``tournament_n``-way tournaments on IBEA fitness over pool + offspring
    (:mod:`selection`; the archive is that whole population, as bluepyopt's
    selIBEA with alpha = its size).  ``lam`` must equal mu.

    ``cxpb``, ``mutpb``, ``eta``: crossover and mutation probabilities and
    the SBX / polynomial-mutation distribution index.
    ``record(gen, nevals, objectives)`` sees each generation's population
    (or pool) objectives.  ``observe(genomes, features)`` sees every
    simulated batch; ``hof``, a
    :class:`selection.ParetoFront`, sees every simulated individual (the
    reference passes ``hof=tools.ParetoFront()``).  ``cp_filename`` /
    ``cp_frequency`` / ``continue_cp`` checkpoint and resume the run as the
    reference's ``opt.run(cp_filename=..., cp_frequency=..., continue_cp=...)``:
    pool or population, objectives, best individual, hall of fame, history,
    generator state and the caller's arrays from ``save_state()`` (restored
    with ``load_state``); a resumed run continues exactly as the
    uninterrupted one would.  The result's best individual is the one of
    least summed objective among all simulated."""
    if selector not in ("sum", "ibea"):
        raise ValueError(f"selector {selector!r}: use 'sum' or 'ibea'")
    if selector == "ibea" and lam != mu:
        raise ValueError("selector 'ibea' makes one offspring per pool member: lam must equal mu")
    m = evaluator.m
    rng = np.random.default_rng(seed)
    lo = np.log10([fp.lo for fp in m.free])
    hi = np.log10([fp.hi for fp in m.free])
    t0 = time.perf_counter()
    best = [None, np.inf]

    def note_best(x, o):
        s = o.sum(axis=1)
        k = int(np.argmin(s))
        if s[k] < best[1]:
            best[0], best[1] = x[k].copy(), float(s[k])

    def evaluate(genomes):
        obj, feat = evaluator.evaluate(10 ** genomes)
        if observe is not None:
            observe(genomes, feat)
        if hof is not None:
            hof.update(genomes, obj)
        note_best(genomes, obj)
        return obj

    start = 1
    if continue_cp:
        if not cp_filename or not os.path.exists(cp_filename):
            raise FileNotFoundError(f"no checkpoint {cp_filename!r} to continue from")
        cp = _load_checkpoint(cp_filename)
        if str(cp["selector"]) != selector:
            raise ValueError(f"checkpoint {cp_filename!r} was written by selector {str(cp['selector'])!r}")
        rng.bit_generator.state = cp["rng_state"]
        pop, obj = cp["pop"], cp["obj"]
        best[0], best[1] = cp["best_x"], float(cp["best_score"])
        hist = [(int(g), float(t), float(b), int(n)) for g, t, b, n in cp["history"]]
        t0 -= hist[-1][1]
        evaluator.nevals = int(cp["nevals"])
        if hof is not None and "hof_x" in cp:
            hof.x, hof.f = cp["hof_x"], cp["hof_f"]
        if load_state is not None:
            load_state(cp["extra"])
        start = int(cp["gen"]) + 1
    else:
        if init is not None:
            if len(init) < 3:
                raise ValueError("init = (genomes, scores, objective rows): the objective rows enter selection")
            pop, obj = np.asarray(init[0]), np.asarray(init[2])
            note_best(pop, obj)
        else:
            pop = rng.uniform(lo, hi, size=(n_init or mu, lo.size))
            obj = evaluate(pop)
        if pop.shape[0] > mu:
            if selector == "ibea":
                keep = np.sort(tournament(rng, ibea_fitness(eps_indicator(obj))[0], mu, tournament_n))
            else:
                keep = np.argsort(obj.sum(axis=1), kind="stable")[:mu]
            pop, obj = pop[keep], obj[keep]
        hist = [(0, time.perf_counter() - t0, best[1], evaluator.nevals)]
        if record is not None:
            record(0, evaluator.nevals, obj)
    for gen in range(start, ngen + 1):
        if selector == "ibea":
            kids = np.vstack([var_and_rows(rng, pop, lo, hi, cxpb, mutpb, eta) for _ in range(oversample)])
        else:
            kids = var_or_rows(rng, pop, lam * oversample, lo, hi, cxpb, mutpb, eta)
        if screen is not None:
            kids = kids[screen(kids)]
        kobj = evaluate(kids)
        allx, allo = np.vstack([pop, kids]), np.vstack([obj, kobj])
        if selector == "ibea":
            keep = tournament(rng, ibea_fitness(eps_indicator(allo))[0], mu, tournament_n)
        else:
            keep = np.argsort(allo.sum(axis=1), kind="stable")[:mu]
        pop, obj = allx[keep], allo[keep]
        hist.append((gen, time.perf_counter() - t0, best[1], evaluator.nevals))
        if record is not None:
            record(gen, evaluator.nevals, obj)
        if log is not None:
            log(f"gen {gen}: best {hist[-1][2]:.4f} after {hist[-1][1]:.1f} s, {evaluator.nevals} evaluations")
        if cp_filename and cp_frequency and gen % cp_frequency == 0:
            _save_checkpoint(cp_filename, selector, gen, rng, pop, obj, best, hist, hof, evaluator.nevals,
                             save_state() if save_state is not None else None)
        if time_limit is not None and hist[-1][1] > time_limit:
            break
        if stop_score is not None and best[1] <= stop_score:
            break
    return GAResult(10 ** best[0], best[1], hist, pop, obj, hof)


class DEAPOptimisation:
    """The interface of bluepyopt.optimisations.DEAPOptimisation as the
    reference's driver uses it (Figures/Figure4_DEAP/optimize_parameters_genetic_alg_par.py:50-55):
    ``DEAPOptimisation(evaluator, seed=, offspring_size=, eta=, mutpb=, cxpb=,
    hof=ParetoFront())`` then ``pop, hof, log, hist = opt.run(max_ngen=,
    cp_filename=, cp_frequency=, continue_cp=)``.  The evaluator is an
    :class:`Evaluator` (simulation and features on the device); the search
    runs in log10 parameter space with the IBEA selector.  ``pop`` holds
    parameter rows, ``hof`` the Pareto front (log10 genomes in ``hof.x``),
    ``log`` one record per generation run by this call (gen, nevals and
    min/avg/max/std of the summed objective over the selected pool),
    ``hist`` (generation, seconds, best score, evaluations) rows since the
    start (a resumed run's come from its checkpoint)."""

    def __init__(self, evaluator, seed: int = 1, offspring_size: int = 10, eta: float = 10.0,
                 mutpb: float = 1.0, cxpb: float = 1.0, hof: Optional[ParetoFront] = None,
                 selector_name: str = "IBEA"):
        if selector_name.upper() != "IBEA":
            raise ValueError(f"selector {selector_name!r}: only IBEA")
        self.evaluator = evaluator
        self.seed, self.offspring_size = seed, offspring_size
        self.eta, self.mutpb, self.cxpb = eta, mutpb, cxpb
        self.hof = hof if hof is not None else ParetoFront()

    def run(self, max_ngen: int = 10, cp_filename: Optional[str] = None, cp_frequency: int = 1,
            continue_cp: bool = False):
        log = []

        def record(gen, nevals, obj):
            t = obj.sum(axis=1)
            log.append({"gen": gen, "nevals": nevals, "min": float(t.min()), "avg": float(t.mean()),
                        "max": float(t.max()), "std": float(t.std())})
        res = run_ga(self.evaluator, mu=self.offspring_size, lam=self.offspring_size, ngen=max_ngen,
                     seed=self.seed, selector="ibea", hof=self.hof, cp_filename=cp_filename,
                     cp_frequency=cp_frequency, continue_cp=continue_cp, cxpb=self.cxpb,
                     mutpb=self.mutpb, eta=self.eta, record=record)
        return 10 ** res.population, self.hof, log, res.history
