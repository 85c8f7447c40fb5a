"""Multi-objective selection for the GA: IBEA and a Pareto-front archive.

The reference optimises with bluepyopt 1.9.126's ``DEAPOptimisation``
(Figures/Figure4_DEAP/optimize_parameters_genetic_alg_par.py:50), whose
default selector is IBEA and whose hall of fame there is deap's
``tools.ParetoFront()``.  Neither library is available, so both are restated
here from the published algorithms:

* IBEA -- Zitzler & Kuenzli, "Indicator-Based Selection in Multiobjective
  Search" (PPSN VIII, 2004), with the additive epsilon indicator on
  objectives normalised by their range: I(a, b) = max_i (f_i(a) - f_i(b)) /
  range_i, the smallest shift that makes a weakly dominate b (objectives
  minimised).  Fitness F(x) = sum_{y != x} -exp(-I(y, x) / (c kappa)) with
  c = max |I| and kappa = 0.05; environmental selection removes the
  individual of least fitness and adds its term back to the others until the
  archive holds alpha individuals; mating selection is a tournament on F.
* Pareto front -- the individuals no other evaluated individual dominates
  (every objective <=, one <).

Both work on objective matrices [n, nobj] (lower is better).
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

KAPPA = 0.05
IBEA_MAX = 8192           # individuals in one environmental selection (two n x n matrices)


def eps_indicator(F: np.ndarray) -> np.ndarray:
    """I[a, b] = max_i (F[a, i] - F[b, i]) / range_i for all pairs (ranges of
    zero count as 1)."""
    F = np.asarray(F, dtype=np.float64)
    rng = F.max(axis=0) - F.min(axis=0)
    rng[rng == 0] = 1.0
    G = F / rng
    n, m = G.shape
    out = np.empty((n, n))
    step = max(1, (1 << 22) // max(1, n * m))
    for a0 in range(0, n, step):
        a1 = min(n, a0 + step)
        out[a0:a1] = np.max(G[a0:a1, None, :] - G[None, :, :], axis=2)
    return out


def ibea_fitness(I: np.ndarray, kappa: float = KAPPA) -> Tuple[np.ndarray, np.ndarray]:
    """Fitness of every individual and the matrix of terms T[y, x] =
    -exp(-I[y, x] / (c kappa)) it sums (diagonal excluded)."""
    c = float(np.max(np.abs(I)))
    if c == 0.0:
        c = 1.0
    T = -np.exp(-I / (c * kappa))
    np.fill_diagonal(T, 0.0)
    return T.sum(axis=0), T


def environmental_selection(F: np.ndarray, alpha: int, kappa: float = KAPPA) -> Tuple[np.ndarray, np.ndarray]:
    """Indices of the alpha individuals IBEA keeps (in their original order)
    and their fitness after the removals."""
    n = F.shape[0]
    if alpha >= n:
        fit, _ = ibea_fitness(eps_indicator(F), kappa)
        return np.arange(n), fit
    if n > IBEA_MAX:
        raise ValueError(f"IBEA on {n} individuals needs an {n} x {n} indicator matrix; at most {IBEA_MAX}")
    fit, T = ibea_fitness(eps_indicator(F), kappa)
    alive = np.ones(n, dtype=bool)
    work = fit.copy()
    for _ in range(n - alpha):
        w = np.where(alive, work, np.inf)
        x = int(np.argmin(w))
        alive[x] = False
        work -= T[x]                  # remove x's term from everyone's sum
    keep = np.nonzero(alive)[0]
    return keep, work[keep]


def tournament(rng, fitness: np.ndarray, k: int, size: int = 2) -> np.ndarray:
    """k winners of tournaments of ``size`` uniformly drawn individuals
    (highest fitness wins)."""
    cand = rng.integers(fitness.size, size=(k, size))
    return cand[np.arange(k), np.argmax(fitness[cand], axis=1)]


def dominated(F: np.ndarray, by: Optional[np.ndarray] = None) -> np.ndarray:
    """mask[i]: some row of ``by`` (default F) dominates F[i]."""
    F = np.asarray(F, dtype=np.float64)
    B = F if by is None else np.asarray(by, dtype=np.float64)
    out = np.zeros(F.shape[0], dtype=bool)
    step = max(1, (1 << 22) // max(1, B.shape[0] * B.shape[1]))
    for i0 in range(0, F.shape[0], step):
        i1 = min(F.shape[0], i0 + step)
        le = np.all(B[None, :, :] <= F[i0:i1, None, :], axis=2)
        lt = np.any(B[None, :, :] < F[i0:i1, None, :], axis=2)
        out[i0:i1] = np.any(le & lt, axis=1)
    return out


class ParetoFront:
    """Every evaluated individual that no other evaluated individual
    dominates (deap.tools.ParetoFront's role in the reference's driver);
    duplicates of an archived objective vector are kept once."""

    def __init__(self):
        self.x = np.zeros((0, 0))
        self.f = np.zeros((0, 0))

    def update(self, x: np.ndarray, f: np.ndarray) -> None:
        x = np.atleast_2d(np.asarray(x, dtype=np.float64))
        f = np.atleast_2d(np.asarray(f, dtype=np.float64))
        if self.f.size == 0:
            X, Fo = x, f
        else:
            X, Fo = np.vstack([self.x, x]), np.vstack([self.f, f])
        keep = ~dominated(Fo)
        X, Fo = X[keep], Fo[keep]
        _, first = np.unique(Fo, axis=0, return_index=True)
        first.sort()
        self.x, self.f = X[first], Fo[first]

    def __len__(self) -> int:
        return self.f.shape[0]
