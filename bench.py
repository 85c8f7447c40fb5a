"""Benchmark of the simulation loop on one MI355X.

One bench step = one pass of the simulation over the whole batch: every
(parameter set, stimulus) run advanced through a window of dt steps by
libneurogpu's HIP kernel, inputs already resident in device memory.  The
headline workload is the thick-tufted-pyramidal-like cell (1,200 compartments,
the 13 BBP TTPC mechanisms), 8,192 parameter sets x 5 step currents,
dt = 0.025 ms; a bench step is its whole 300 ms protocol, 12,000 dt steps in
one persistent launch.  ``--workload active500`` measures the active
500-compartment cell over its full 10,000 steps; ``--nt`` shortens the
protocol (profiling runs).

Prints one JSON line (the last line of stdout).
"""
import argparse
import json
import os
import sys
import time

_ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, _ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
METRIC = "compartment-steps/s at FP64 oracle <=1e-6 mV"


def algorithmic_bytes_per_comp_step(model) -> float:
    """B_cs = 16 (1 + S) + 8 N_rec / N_c: every state variable and v read and
    written once per compartment-step, plus the recorded trace."""
    return 16.0 * (1.0 + model.state_count()) + 8.0 * len(model.rec_sites) / model.tree.N


def traffic_from_profile(name: str):
    """HBM bytes per launch recorded by the rocprofv3 PMC passes, if present
    (profiles/<name>_traffic.txt, line 'traffic_bytes_per_launch: <value>')."""
    path = os.path.join(_ROOT, "profiles", f"{name}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def fp64_from_profile(name: str):
    """Issued FP64 flops per compartment-step from the rocprofv3 FP64
    instruction-count pass (profiles/<name>_fp64.txt, line
    'fp64_flop_per_comp_step: <value>'), if present."""
    path = os.path.join(_ROOT, "profiles", f"{name}_fp64.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("fp64_flop_per_comp_step:"):
            return float(line.split(":", 1)[1])
    return None


def copy_peak_gbs(mib: int = 2048, reps: int = 10) -> float:
    """Measured device-to-device copy bandwidth (read + write bytes / time), in
    a child process: torch ships its own HIP runtime, which must not share a
    process with the one libneurogpu links."""
    import subprocess
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-peak", str(mib), "--steps", str(reps)],
                         capture_output=True, text=True, timeout=300)
    if out.returncode != 0:
        raise RuntimeError(f"copy peak measurement failed: {out.stderr[-2000:]}")
    return float(json.loads(out.stdout.strip().splitlines()[-1])["copy_peak_GBs"])


def _copy_peak_child(mib: int, reps: int) -> None:
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    e.synchronize()
    print(json.dumps({"copy_peak_GBs": 2.0 * n * 8 * reps / (s.elapsed_time(e) * 1e-3) / 1e9}))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, nsets: int, threads: int):
    """The FP64 oracle on a bounded sample of the same workload."""
    from oracle import cpu_ref
    m = w.model
    x = w.x[:nsets]
    t0 = time.perf_counter()
    ref = cpu_ref.simulate(m.tree, m.mechset_id, m.allparams(x), w.amps, w.dts, m.stim_seg, m.stim_area,
                           m.v_init, m.rec_sites, nthreads=threads)
    el = time.perf_counter() - t0
    cs = nsets * w.amps.shape[0] * m.tree.N * w.dts.size
    return ref, cs / el, el


def spike_counts(tr, thr: float = -30.0):
    """Upward crossings of thr along the last axis."""
    return np.sum((tr[..., 1:] >= thr) & (tr[..., :-1] < thr), axis=-1)


def bench_mixed(args):
    """The uneven batch: four morphologies of 200 / 700 / 1,500 / 3,000
    segments, 12,000 / 4,000 / 1,000 / 256 instances, 10,000 / 5,000 / 2,000 /
    40,000 steps, in one persistent launch through the device work queue; the
    launch shape of each morphology comes from the autotuner (untimed)."""
    from neurogpu_amd import autotune, models
    from neurogpu_amd._lib import Batch, workload_simulator
    ws = models.mixed_workloads()
    say = lambda m: print(m, flush=True)
    choices = {}
    for w in ws:
        say(f"tuning {w.name}: N={w.model.tree.N} runs={w.nruns} nt={w.dts.size}")
        c, _ = autotune.tune(w, log=say)
        choices[w.name] = c
    sims = [workload_simulator(w) for w in ws]
    b = Batch(0)
    for i, (w, sim) in enumerate(zip(ws, sims)):
        b.add(sim)
        b.set_launch(i, **choices[w.name].as_dict())
    for i in range(args.warmup):
        st = b.run()
        say(f"warmup {i}: kernel {st['ms_kernel']:.1f} ms")
    t0 = time.perf_counter()
    kms = []
    for i in range(args.steps):
        st = b.run()
        kms.append(st["ms_kernel"])
        say(f"step {i}: kernel {st['ms_kernel']:.1f} ms")
    wall = time.perf_counter() - t0
    infos = [b.member_info(i) for i in range(len(ws))]
    total = sum(w.comp_steps for w in ws)
    launch_ms = float(np.mean(kms))
    alg = sum(w.comp_steps * algorithmic_bytes_per_comp_step(w.model) for w in ws)
    achieved = alg / (launch_ms * 1e-3) / 1e9
    per = []
    for w, info in zip(ws, infos):
        span = info["last_end_ms"] - info["first_start_ms"]
        per.append({"morphology": w.name, "segments": w.model.tree.N, "instances": int(w.x.shape[0]),
                    "steps": int(w.dts.size), "comp_steps": w.comp_steps,
                    "first_start_ms": round(info["first_start_ms"], 2), "last_end_ms": round(info["last_end_ms"], 2),
                    "comp_steps_per_s_in_span": w.comp_steps / (span * 1e-3) if span > 0 else None,
                    "launch": {k: info[k] for k in ("waves_per_run", "runs_per_group", "runs_per_item", "block_sync",
                                                    "lds_staging")}})
    result = {
        "metric": METRIC, "value": total * args.steps / wall, "unit": "comp-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "mixed4: TTPC mechanisms on 200/700/1500/3000-segment cells, 12000/4000/1000/256 "
                               "instances, 10000/5000/2000/40000 steps, one persistent launch", "items": st["nitems"],
                   "workgroups": st["workgroups"], "lds_bytes": st["lds_bytes"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms_per_launch": round(launch_ms, 3)},
        "per_morphology": per,
        "tail_ms": round(max(i["last_end_ms"] for i in infos) - sorted(i["last_end_ms"] for i in infos)[-2], 2),
    }
    b.close()
    if args.separate:
        sep = {}
        for w, sim in zip(ws, sims):
            _, st1 = sim.run(fetch=False)
            _, st1 = sim.run(fetch=False)
            sep[w.name] = st1["ms_kernel"]
            say(f"separate {w.name}: {st1['ms_kernel']:.1f} ms")
        result["separate_launches_ms"] = sep
        result["separate_total_ms"] = sum(sep.values())
    for sim in sims:
        sim.close()
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ttpc1200", choices=("passive64", "active500", "ttpc1200", "mixed4"))
    ap.add_argument("--separate", action="store_true",
                    help="mixed4: also time each morphology in its own launch, one after another")
    ap.add_argument("--nsets", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--waves-per-run", type=int, default=0)
    ap.add_argument("--runs-per-group", type=int, default=0)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--cpu-sets", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy-peak", action="store_true")
    ap.add_argument("--copy-peak", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_peak:
        _copy_peak_child(args.copy_peak, args.steps)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (runs are independent, "
              f"so the throughput of N GPUs is N times this figure).", file=sys.stderr)
        sys.exit(2)

    from neurogpu_amd import models
    from neurogpu_amd._lib import Simulator
    if args.workload == "mixed4":
        return bench_mixed(args)

    w = models.workload(args.workload, nsets=args.nsets, nt=args.nt)
    m = w.model
    print(f"bench: {w.name} N={m.tree.N} runs={w.nruns} nt={w.dts.size} S={m.state_count():.2f}", flush=True)
    sim = Simulator(0)
    sim.set_tree(m.tree, m.mechset_id)
    sim.set_sim(m.v_init, m.rec_sites)
    sim.set_params(w.params, m.ncomps, m.nparams)
    sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
    sim.set_launch(args.waves_per_run, args.runs_per_group, 0, bool(args.graph))
    for i in range(args.warmup):
        _, st = sim.run(fetch=False)
        print(f"warmup {i}: kernel {st['ms_kernel']:.1f} ms", flush=True)
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = sim.run(fetch=False)
        kernel_ms.append(st["ms_kernel"])
        print(f"step {i}: kernel {st['ms_kernel']:.1f} ms", flush=True)
    wall = time.perf_counter() - t0
    gpu = sim.fetch().reshape(w.x.shape[0], w.amps.shape[0], len(m.rec_sites), w.dts.size)
    sim.close()

    comp_steps = w.comp_steps
    value = comp_steps * args.steps / wall
    launch_ms = float(np.mean(kernel_ms)) / max(1, st["launches"])
    bcs = algorithmic_bytes_per_comp_step(m)
    alg_bytes_per_launch = comp_steps / max(1, st["launches"]) * bcs
    achieved = alg_bytes_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = traffic_from_profile(args.workload)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_comp_step": round(bcs, 3), "kernel_ms_per_launch": round(launch_ms, 3),
                "comp_steps_per_launch": comp_steps / max(1, st["launches"])}
    if traffic:
        roofline["actual_GBs"] = round(traffic / (launch_ms * 1e-3) / 1e9, 1)
    if not args.no_copy_peak:
        roofline["copy_peak_GBs"] = round(copy_peak_gbs(), 1)
    flop_cs = fp64_from_profile(args.workload)
    if flop_cs:
        tf = flop_cs * comp_steps / max(1, st["launches"]) / (launch_ms * 1e-3) / 1e12
        roofline["fp64"] = {"issued_flop_per_comp_step": flop_cs, "achieved": round(tf, 2),
                            "peak": FP64_PEAK_TFS, "unit": "TFLOP/s", "frac": round(tf / FP64_PEAK_TFS, 4)}

    result = {
        "metric": METRIC, "value": value, "unit": "comp-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.name}: {m.tree.N} compartments, {m.mechset} mechanisms, "
                               f"{w.x.shape[0]} sets x {w.amps.shape[0]} stimuli, dt {w.dts[0]} ms, "
                               f"{w.dts.size} steps", "runs": w.nruns, "compartments": m.tree.N,
                   "nt": int(w.dts.size), "launch": {k: st[k] for k in ("waves_per_run", "runs_per_group", "slots_per_lane",
                                                                      "runs_per_block", "launches")}},
        "roofline": roofline,
    }
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        cpu_sets = args.cpu_sets if args.cpu_sets is not None else (16 if args.workload == "ttpc1200" else 8)
        nsets = min(cpu_sets, w.x.shape[0])
        ref, cps, el = cpu_baseline(w, nsets, threads)
        d = float(np.max(np.abs(gpu[:nsets] - ref)))
        same_spikes = bool(np.array_equal(spike_counts(gpu[:nsets]), spike_counts(ref)))
        one_sets = max(1, nsets // threads)
        _, cps1, el1 = cpu_baseline(w, one_sets, 1)
        result["cpu_baseline"] = {"value": cps, "unit": "comp-steps/s", "cores": threads, "kind": "port",
                                  "sample": f"{nsets} sets x {w.amps.shape[0]} stimuli x {w.dts.size} steps "
                                            f"of the same workload ({el:.1f} s, OpenMP over runs)",
                                  "cpu": cpu_model(), "nproc": os.cpu_count(),
                                  "one_thread": {"value": cps1, "sample": f"{one_sets} sets ({el1:.1f} s)"}}
        result["parity_sample_max_abs_mV"] = d
        result["parity_sample_spike_counts_equal"] = same_spikes
        if d > 1e-6 or not same_spikes:
            print(f"PARITY FAILURE on the sample: max |dv| {d} mV, spike counts equal: {same_spikes}",
                  file=sys.stderr)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
