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
bluepyopt's DEAP setup uses; selection keeps the mu best by total objective
(the reference's IBEA selector is not reproduced).
"""
from __future__ import annotations

import dataclasses
import time
from typing import Callable, List, Optional

import numpy as np

from ._lib import FEATURES, Simulator

PENALTY = 250.0          # objective value for a feature that is undefined (efel_ext.py:73-83 uses 10000 / 100000)
GA_FEATURES = (0, 1, 2, 4)   # spike count, first-spike latency, AP amplitude, mean ISI per stimulus


def fill_features(f: np.ndarray, window: float) -> np.ndarray:
    """Features with the undefined entries replaced by their natural limits
    (no spike: latency and ISI = the stimulus window, amplitude = width = 0),
    so that they can be regressed and compared."""
    out = np.array(f, dtype=np.float64, copy=True)
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


def objectives(filled: np.ndarray, target: Target) -> np.ndarray:
    """|f - f*| / scale per stimulus and GA feature -> [n, nstim * len(GA_FEATURES)]."""
    d = np.abs(filled[..., GA_FEATURES] - target.features[None, :, GA_FEATURES]) / target.scale[list(GA_FEATURES)]
    d = np.where(np.isfinite(d), d, PENALTY)
    return np.minimum(d, PENALTY).reshape(filled.shape[0], -1)


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
        return fill_features(f, self.window)

    def evaluate(self, x: np.ndarray):
        f = self.features(x)
        return objectives(f, self.target), f

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


def sbx_rows(A, B, lo, hi, u_gene, u_beta, u_swap, eta=20.0):
    """First children of bounded SBX for rows of parents A, B [n, g], from
    given uniform numbers [n, g] (gene crossed when u_gene <= 0.5, spread
    from u_beta, order from u_swap): the same arithmetic as :func:`sbx`, one
    array operation per step instead of one Python step per gene."""
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
    return np.where(cross, np.where(u_swap <= 0.5, y2, y1), A)


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


def run_ga(evaluator: Evaluator, mu: int, lam: int, ngen: int, seed: int = 1178,
           screen: Optional[Callable] = None, oversample: int = 1, time_limit: Optional[float] = None,
           stop_score: Optional[float] = None, log: Optional[Callable] = None,
           observe: Optional[Callable] = None, n_init: Optional[int] = None,
           init: Optional[tuple] = None) -> GAResult:
    """(mu + lambda) GA in log10 parameter space.  With ``screen`` the GA draws
    ``oversample * lam`` offspring, and ``screen(genomes)`` returns the indices
    of those to simulate (the surrogate pre-screen); every score that enters
    selection comes from the simulator.  ``observe(genomes, features)`` sees
    every simulated batch (the screened GA's surrogate learns from them).
    ``n_init`` random individuals (default ``mu``) are simulated first and
    the best ``mu`` of them form the first population; ``init`` = (genomes,
    simulated scores) gives the first population instead (individuals already
    simulated, e.g. the surrogate's training set)."""
    m = evaluator.m
    rng = np.random.default_rng(seed)
    lo = np.log10([fp.lo for fp in m.free])
    hi = np.log10([fp.hi for fp in m.free])
    t0 = time.perf_counter()
    if init is not None:
        pop, score = np.asarray(init[0]), np.asarray(init[1])
    else:
        pop = rng.uniform(lo, hi, size=(n_init or mu, lo.size))
        obj, feat = evaluator.evaluate(10 ** pop)
        if observe is not None:
            observe(pop, feat)
        score = obj.sum(axis=1)
    first = np.argsort(score, kind="stable")[:mu]
    pop, score = pop[first], score[first]
    hist = [(0, time.perf_counter() - t0, float(score.min()), evaluator.nevals)]
    for gen in range(1, ngen + 1):
        kids = var_or_rows(rng, pop, lam * oversample, lo, hi)
        if screen is not None:
            kids = kids[screen(kids)]
        kobj, kfeat = evaluator.evaluate(10 ** kids)
        if observe is not None:
            observe(kids, kfeat)
        allx = np.vstack([pop, kids])
        alls = np.concatenate([score, kobj.sum(axis=1)])
        keep = np.argsort(alls, kind="stable")[:mu]
        pop, score = allx[keep], alls[keep]
        hist.append((gen, time.perf_counter() - t0, float(score.min()), evaluator.nevals))
        if log is not None:
            log(f"gen {gen}: best {hist[-1][2]:.4f} after {hist[-1][1]:.1f} s, {evaluator.nevals} evaluations")
        if time_limit is not None and hist[-1][1] > time_limit:
            break
        if stop_score is not None and score.min() <= stop_score:
            break
    best = int(np.argmin(score))
    return GAResult(10 ** pop[best], float(score[best]), hist)
