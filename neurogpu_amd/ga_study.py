"""Plain GA vs surrogate-screened GA on one GPU, both scored by simulation.

    python -m neurogpu_amd.ga_study [--workload ttpc1200] [--pop 1024] [--gens 25] ...

The plain (mu + lambda) GA runs ``--gens`` generations; its best simulated
score S* and wall time T are recorded.  The screened GA first simulates
``--train`` log-uniform samples, trains the MLP ensemble on their features
(both timed), then runs generations in which 10x oversampled offspring are
ranked by the surrogate and only the top 10 % plus the most uncertain 2 % are
simulated, until its best simulated score reaches S*.  Prints one JSON line
with both histories, the time to reach S*, and evaluations per second.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import ga, models
from .surrogate import MLPEnsemble, feature_targets, genome_to_input, make_screen


def study(name="ttpc1200", pop=1024, gens=25, train=8192, nt=2000, members=5, epochs=60, seed=1178, device=0,
          max_screen_gens=None):
    w = models.workload(name, nsets=1, nt=nt)
    m = w.model
    dt = float(w.dts[0])
    t_on, t_off = 5.0, 45.0 if nt * dt >= 45.0 else nt * dt
    amps = {"active500": [0.2, 0.5, 1.0], "ttpc1200": [0.1, 0.3, 0.5, 0.7, 0.9]}.get(name, [0.5])
    w.amps, w.dts = models.step_stimulus(amps, nt, dt, t_on, t_off)
    nstim = len(amps)
    ev = ga.Evaluator(w, t_on, t_off, device=device, target_x=m.nominal())
    lo = np.log10([fp.lo for fp in m.free])
    hi = np.log10([fp.hi for fp in m.free])

    ev.nevals = 0
    say = lambda msg: print(msg, flush=True)
    say(f"plain GA: {w.name}, pop {pop}, {gens} generations")
    plain = ga.run_ga(ev, mu=pop, lam=pop, ngen=gens, seed=seed, log=say)
    plain_t = plain.history[-1][1]
    s_star = plain.best_score
    plain_evals = ev.nevals

    ev.nevals = 0
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed + 1)
    g = rng.uniform(lo, hi, size=(train, lo.size))
    f = ev.features(10 ** g)
    say(f"surrogate: {train} training simulations in {time.perf_counter() - t0:.1f} s")
    ens = MLPEnsemble(lo.size, nstim * len(ga.GA_FEATURES), hidden=256, members=members, seed=seed, device=device)
    losses = ens.fit(genome_to_input(g, lo, hi), feature_targets(f), epochs=epochs, batch=512, lr=2e-3)
    prep_t = time.perf_counter() - t0
    say(f"surrogate: trained {members} members, final loss {losses[:, -1].mean():.4f}, prep {prep_t:.1f} s")
    screen = make_screen(ens, ev.target, lo, hi, nstim, keep_frac=0.1, disagree_frac=0.02)
    remaining = None if plain_t is None else max(1.0, 4 * plain_t - prep_t)
    scr = ga.run_ga(ev, mu=pop, lam=pop, ngen=max_screen_gens or 10 * gens, seed=seed, screen=screen, oversample=10,
                    stop_score=s_star, time_limit=remaining, log=say)
    scr_hist = [(gen, t + prep_t, s, n) for gen, t, s, n in scr.history]
    reached = next((t for _, t, s, _ in scr_hist if s <= s_star), None)
    ens.close()
    ev.close()
    return {
        "workload": f"{w.name} protocol {nt} steps x {nstim} stimuli, pop {pop}",
        "plain": {"best_score": s_star, "wall_s": plain_t, "evals": plain_evals,
                  "evals_per_s": plain_evals / plain_t, "history": plain.history},
        "screened": {"best_score": scr.best_score, "prep_s": prep_t, "train_samples": train,
                     "final_loss": float(losses[:, -1].mean()), "evals": ev.nevals,
                     "history": scr_hist, "wall_s_to_reach_plain_best": reached,
                     "evals_per_s": ev.nevals / scr_hist[-1][1]},
        "time_ratio": (reached / plain_t) if reached else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ttpc1200", choices=models.WORKLOADS)
    ap.add_argument("--pop", type=int, default=1024)
    ap.add_argument("--gens", type=int, default=25)
    ap.add_argument("--train", type=int, default=8192)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--members", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=60)
    args = ap.parse_args()
    r = study(args.workload, args.pop, args.gens, args.train, args.nt, args.members, args.epochs)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
