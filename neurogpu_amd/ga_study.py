"""Plain GA vs surrogate-screened GA on one GPU, both scored by simulation.

    python -m neurogpu_amd.ga_study [--workload ttpc1200] [--pop 1024] [--gens 20] ...

The plain (mu + lambda) GA runs ``--gens`` generations; its best simulated
score S* and wall time T are recorded.  The screened GA starts from the same
simulated random population, trains the MLP ensemble on it, then runs
generations in which ``--oversample``-fold offspring are ranked by the
surrogate and only the best ``--keep`` of a generation's worth plus the most
uncertain ``--disagree`` (the disagreement gate) are simulated.  Every
simulated batch is added to the training set and the ensemble continues
training on it (``--refit-epochs`` per generation), so the surrogate learns
where the search goes.  The run stops when its best simulated score reaches
S*; all of its time (training included) counts.  ``--pretrain N`` first
simulates N seeded log-uniform individuals and trains the ensemble on them
(the offline training set of the study at population 16,384); that time
counts as well, and the screened GA then starts from the best ``--pop`` of
those simulated individuals instead of simulating a population of its own.  Prints one JSON line with both histories, the time to reach
S*, and evaluations per second.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import ga, models
from .surrogate import MLPEnsemble, feature_targets, genome_to_input, make_screen


def study(name="ttpc1200", pop=1024, gens=20, nt=2000, members=5, epochs=60, seed=1178, device=0,
          oversample=20, keep=0.15, disagree=0.03, refit_epochs=10, explore=0.0, max_train=4096,
          max_screen_gens=None, mu_screen=None, random=0.0, pretrain=0, protocol="short"):
    """``protocol="short"``: ``nt`` steps with the step current on from 5 to
    45 ms; ``"full"``: the workload's own protocol (active500: 10,000 steps
    of 0.02 ms, 20-180 ms), ``nt`` ignored.  Features are measured while the
    current is on."""
    if protocol == "full":
        w = models.workload(name, nsets=1)
        nt = int(w.dts.size)
        t_on, t_off = models.STIM_WINDOW[name]
    else:
        w = models.workload(name, nsets=1, nt=nt)
        dt = float(w.dts[0])
        t_on, t_off = 5.0, 45.0 if nt * dt >= 45.0 else nt * dt
        amps = {"active500": [0.2, 0.5, 1.0], "ttpc1200": [0.1, 0.3, 0.5, 0.7, 0.9]}.get(name, [0.5])
        w.amps, w.dts = models.step_stimulus(amps, nt, dt, t_on, t_off)
    m = w.model
    nstim = w.amps.shape[0]
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
    ens = MLPEnsemble(lo.size, nstim * len(ga.GA_FEATURES), hidden=256, members=members, seed=seed, device=device)
    data_g, data_y = [], []
    fit = {"s": 0.0, "n": 0, "loss": float("nan")}
    t_pre = time.perf_counter()
    init = None
    if pretrain > 0:
        # offline training set: seeded log-uniform samples, simulated in batches
        # of one population; its time counts towards the screened GA's
        prng = np.random.default_rng(seed + 7919)
        pre_g, pre_s, pre_o = [], [], []
        for b0 in range(0, pretrain, pop):
            g = prng.uniform(lo, hi, size=(min(pop, pretrain - b0), lo.size))
            obj, feats = ev.evaluate(10 ** g)
            data_g.append(genome_to_input(g, lo, hi))
            data_y.append(feature_targets(feats))
            pre_g.append(g)
            pre_s.append(obj.sum(axis=1))
            pre_o.append(obj)
            say(f"pretraining set: {b0 + g.shape[0]} of {pretrain} simulated")
        # the screened GA starts from the best of these simulated individuals
        init = (np.vstack(pre_g), np.concatenate(pre_s), np.vstack(pre_o))
        t = time.perf_counter()
        losses = ens.fit(np.vstack(data_g), np.vstack(data_y), epochs=epochs, batch=512, lr=2e-3)
        fit.update(s=time.perf_counter() - t, n=1, loss=float(losses[:, -1].mean()))
        say(f"surrogate trained on {pretrain} samples in {fit['s']:.1f} s, loss {fit['loss']:.4f}")
    pre_t = time.perf_counter() - t_pre

    def observe(genomes, feats):
        """Every simulated batch joins the training set; the ensemble trains
        from scratch on the first (the random population) and continues on
        the newest ``max_train`` samples after each generation."""
        t = time.perf_counter()
        data_g.append(genome_to_input(genomes, lo, hi))
        data_y.append(feature_targets(feats))
        x, y = np.vstack(data_g)[-max_train:], np.vstack(data_y)[-max_train:]
        if fit["n"] == 0:
            losses = ens.fit(x, y, epochs=epochs, batch=512, lr=2e-3)
        else:
            losses = ens.refit(x, y, epochs=refit_epochs, batch=512, lr=1e-3)
        fit["loss"] = float(losses[:, -1].mean())
        fit["n"] += 1
        fit["s"] += time.perf_counter() - t

    screen = make_screen(ens, ev.target, lo, hi, nstim, keep_frac=keep / oversample,
                         disagree_frac=disagree / oversample, explore=explore, random_frac=random / oversample,
                         seed=seed)
    say(f"screened GA: oversample {oversample}, simulate the best {keep:.0%} + most uncertain {disagree:.0%} "
        f"of a generation, surrogate retrained every generation")
    scr = ga.run_ga(ev, mu=mu_screen or pop, lam=pop, ngen=max_screen_gens or 10 * gens, seed=seed, screen=screen,
                    oversample=oversample, stop_score=s_star, time_limit=max(plain_t - pre_t, 0.0), log=say,
                    observe=observe, init=init,
                    n_init=pop)
    scr_hist = [(g, t + pre_t, b, n) for g, t, b, n in scr.history]
    reached = next((t for _, t, s, _ in scr_hist if s <= s_star), None)
    ens.close()
    ev.close()
    return {
        "workload": f"{w.name} protocol {nt} steps x {nstim} stimuli, current {t_on}-{t_off} ms, pop {pop}",
        "plain": {"best_score": s_star, "wall_s": plain_t, "evals": plain_evals,
                  "evals_per_s": plain_evals / plain_t, "history": plain.history},
        "screened": {"best_score": scr.best_score, "evals": ev.nevals, "surrogate_fit_s": fit["s"],
                     "surrogate_fits": fit["n"], "final_loss": fit["loss"], "oversample": oversample,
                     "pretrain": pretrain, "pretrain_s": pre_t,
                     "keep": keep, "disagree": disagree, "explore": explore, "random": random,
                     "mu": mu_screen or pop,
                     "history": scr_hist,
                     "wall_s_to_reach_plain_best": reached, "evals_per_s": ev.nevals / scr_hist[-1][1]},
        "time_ratio": (reached / plain_t) if reached else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ttpc1200", choices=models.WORKLOADS)
    ap.add_argument("--pop", type=int, default=1024)
    ap.add_argument("--gens", type=int, default=20)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--protocol", default="short", choices=("short", "full"),
                    help="full: the workload's own protocol (active500: 10,000 steps, current 20-180 ms)")
    ap.add_argument("--members", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--oversample", type=int, default=20)
    ap.add_argument("--keep", type=float, default=0.15)
    ap.add_argument("--disagree", type=float, default=0.03)
    ap.add_argument("--refit-epochs", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1178)
    ap.add_argument("--explore", type=float, default=0.0)
    ap.add_argument("--max-train", type=int, default=4096)
    ap.add_argument("--mu-screen", type=int, default=None, help="parents of the screened GA (default --pop)")
    ap.add_argument("--pretrain", type=int, default=0,
                    help="random individuals simulated to train the surrogate before the screened GA (time counted)")
    ap.add_argument("--random", type=float, default=0.0,
                    help="fraction of a generation simulated unscreened (uniform over the candidates)")
    args = ap.parse_args()
    r = study(args.workload, args.pop, args.gens, args.nt, args.members, args.epochs, seed=args.seed,
              oversample=args.oversample, keep=args.keep, disagree=args.disagree, refit_epochs=args.refit_epochs,
              explore=args.explore, max_train=args.max_train, mu_screen=args.mu_screen, random=args.random,
              pretrain=args.pretrain, protocol=args.protocol)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
