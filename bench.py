"""Benchmark of the simulation loop on one MI355X.

One bench step = one pass of the simulation over the whole batch: every
(parameter set, stimulus) run advanced through a window of dt steps by one
launch of libneurogpu's HIP kernel (ngpu_run_steps), inputs already resident
in device memory, the state carried from window to window.  The headline
workload is the thick-tufted-pyramidal-like cell (1,200 compartments, the 13
BBP TTPC mechanisms), 8,192 parameter sets x 5 step currents, dt = 0.025 ms,
300 ms = 12,000 dt steps: ``--steps K`` timed windows of 12,000 / K steps
cover the protocol exactly once, after ``--warmup W`` untimed windows.
``--workload active500`` measures the active 500-compartment cell over its
10,000 steps; ``--nt`` shortens the protocol (profiling runs).

HBM traffic per launch comes from two rocprofv3 PMC passes (FETCH_SIZE,
WRITE_SIZE) over a child process, each calibrated on a device-to-device copy.

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


SIM_KERNEL = "ngpu_sim_kernel"
BATCH_KERNEL = "ngpu_batch_kernel"
ISSUE_COUNTERS = ["SQ_WAVES", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU",
                  "SQ_WAIT_ANY", "SQ_INSTS_LDS", "SQ_INSTS_VALU_FMA_F64"]
COPY_KERNEL = "copyBuffer"


def _counter_rows(outdir: str):
    """Rows of every rocprofv3 counter-collection CSV written under outdir."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _pmc_pass(counter, child_args, kernel: str, timeout: int = 180):
    """One rocprofv3 PMC pass over a child bench process: the counter's value
    (summed over the counter's instances) for each dispatch of kernels whose
    name contains ``kernel``, in dispatch order; for a list of counters (one
    pass collecting them together), a dict of such lists.  None if rocprofv3
    is absent or the pass fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    names = [counter] if isinstance(counter, str) else list(counter)
    with tempfile.TemporaryDirectory(prefix="ngpu_pmc_") as d:
        cmd = [prof, "--pmc"] + names + ["--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__)] + list(child_args)
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        except subprocess.TimeoutExpired:
            return None
        rows = _counter_rows(d)
        if r.returncode != 0 and not rows:
            return None
        per = {n: {} for n in names}
        for row in rows:
            keys = {k.lower(): k for k in row}
            name = row.get(keys.get("kernel_name", ""), "")
            ctr = row.get(keys.get("counter_name", ""), "")
            if kernel not in name or ctr not in per:
                continue
            did = int(row.get(keys.get("dispatch_id", ""), 0) or 0)
            per[ctr][did] = per[ctr].get(did, 0.0) + float(row[keys["counter_value"]])
        out = {n: [v[k] for k in sorted(v)] for n, v in per.items()}
        if isinstance(counter, str):
            return out[counter] or None
        return out if all(out.values()) else None


def calibrated_traffic(child, kernel: str):
    """HBM bytes of the last dispatch of ``kernel`` in a child bench process,
    from FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch), each scaled by
    what it reads for a 1 GiB device-to-device copy (bytes copied / counter
    bytes)."""
    copy = ["--copy-probe", "1024"]
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        sim = _pmc_pass(ctr, child, kernel)
        cal = _pmc_pass(ctr, copy, COPY_KERNEL)
        if not sim or not cal:
            return None
        copied = 1024.0 * 1024.0                      # KiB moved each way by one copy dispatch
        factor = copied / (sum(cal) / len(cal))
        out[ctr] = {"kib_per_dispatch": sim[-1], "calibration_factor": round(factor, 4),
                    "bytes": sim[-1] * 1024.0 * factor}
    out["bytes_per_launch"] = out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]
    return out


def measured_traffic(workload: str, nsets, nt, window: int, solve: str = "auto", generated: bool = False):
    """HBM bytes of one window launch of the simulation kernel, from two
    rocprofv3 PMC passes (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) over a
    child that runs the first two windows, read on the second window's
    dispatch; each counter is scaled by what it reads for a 1 GiB
    device-to-device copy (bytes copied / counter bytes)."""
    child = ["--pmc-child", "--workload", workload, "--window", str(window), "--solve", solve]
    if generated:
        child.append("--generated")
    if nsets is not None:
        child += ["--nsets", str(nsets)]
    if nt is not None:
        child += ["--nt", str(nt)]
    return calibrated_traffic(child, SIM_KERNEL)


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
    ws = models.mixed_workloads(mechset=args.mixed_set)
    say = lambda m: print(m, flush=True)
    if args.mixed_choices:
        choices = {k: autotune.Choice(**v) for k, v in json.loads(args.mixed_choices).items()}
    else:
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
    if args.pmc_child:
        b.run()
        b.close()
        for sim in sims:
            sim.close()
        return
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
        "config": {"workload": f"mixed4: {args.mixed_set} mechanisms on 200/700/1500/3000-segment cells, "
                               "12000/4000/1000/256 instances, 10000/5000/2000/40000 steps, one persistent launch", "items": st["nitems"],
                   "workgroups": st["workgroups"], "lds_bytes": st["lds_bytes"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms_per_launch": round(launch_ms, 3)},
        "per_morphology": per,
        "tail_ms": round(max(i["last_end_ms"] for i in infos) - sorted(i["last_end_ms"] for i in infos)[-2], 2),
    }
    b.close()
    if not args.no_pmc:
        say("rocprofv3 passes over one batch launch (child processes)")
        child = ["--pmc-child", "--workload", "mixed4", "--mixed-set", args.mixed_set, "--mixed-choices",
                 json.dumps({k: c.as_dict() for k, c in choices.items()})]
        traffic = calibrated_traffic(child, BATCH_KERNEL)
        if traffic:
            result["roofline"]["traffic"] = traffic["bytes_per_launch"]
            result["roofline"]["traffic_over_algorithmic"] = round(traffic["bytes_per_launch"] / alg, 3)
            result["roofline"]["actual_GBs"] = round(traffic["bytes_per_launch"] / (launch_ms * 1e-3) / 1e9, 1)
            result["roofline"]["pmc"] = {k: traffic[k] for k in ("FETCH_SIZE", "WRITE_SIZE")}
        issue = _pmc_pass(ISSUE_COUNTERS, child, BATCH_KERNEL, timeout=300)
        if issue:
            c = {k: v[-1] for k, v in issue.items()}
            result["issue"] = {"counters": c, "valu_lane_instructions_per_comp_step": c["SQ_INSTS_VALU"] * 64 / total,
                               "fraction_of_wave_cycles_waiting": c["SQ_WAIT_ANY"] / c["SQ_WAVE_CYCLES"],
                               "fraction_of_wave_cycles_valu_active": c["SQ_ACTIVE_INST_VALU"] / c["SQ_WAVE_CYCLES"]}
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


def cpu_threads() -> int:
    """Host threads of this process's CPU share: the affinity mask, bounded by
    OMP_NUM_THREADS where the environment sets it."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else max(1, n)


def windows(nt: int, k: int):
    """k consecutive step windows [b_i, b_i+1) that cover 0..nt once."""
    b = np.linspace(0, nt, k + 1).round().astype(int)
    return [(int(b[i]), int(b[i + 1])) for i in range(k) if b[i + 1] > b[i]]


GENERATED_SET = {"ttpc": 3, "mainen": 4}        # NGPU_MECH_TTPC_GEN, NGPU_MECH_MAINEN_GEN


def mechset_of(w, args) -> int:
    if getattr(args, "generated", False):
        if w.model.mechset not in GENERATED_SET:
            raise SystemExit(f"no generated mechanism set for {w.model.mechset}")
        return GENERATED_SET[w.model.mechset]
    return w.model.mechset_id


def configured_simulator(w, args):
    from neurogpu_amd._lib import Simulator
    m = w.model
    sim = Simulator(0)
    sim.set_tree(m.tree, mechset_of(w, args))
    sim.set_sim(m.v_init, m.rec_sites)
    sim.set_params(w.params, m.ncomps, m.nparams)
    sim.set_stimuli(w.amps, w.dts, m.stim_seg, m.stim_area)
    sim.set_launch(args.waves_per_run, args.runs_per_group, 0, False)
    sim.set_solve({"auto": 0, "reference": 1, "fast": 2}[args.solve])
    return sim


def pmc_child(args):
    """The profiled child of a PMC pass: the first two windows of the workload."""
    from neurogpu_amd import models
    w = models.workload(args.workload, nsets=args.nsets, nt=args.nt)
    sim = configured_simulator(w, args)
    n = min(args.window, w.dts.size)
    sim.run_steps(0, n)
    if 2 * n <= w.dts.size:
        sim.run_steps(n, 2 * n)
    sim.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed bench steps: windows of nt/steps dt steps that cover the protocol once")
    ap.add_argument("--warmup", type=int, default=5, help="untimed windows before the timed protocol")
    ap.add_argument("--workload", default="ttpc1200", choices=("passive64", "active500", "ttpc1200", "mixed4"))
    ap.add_argument("--separate", action="store_true",
                    help="mixed4: also time each morphology in its own launch, one after another")
    ap.add_argument("--mixed-set", default="mainen", choices=("mainen", "ttpc"),
                    help="mixed4: the channel set of the four cells (mainen = the active cell's)")
    ap.add_argument("--mixed-choices", default="", help=argparse.SUPPRESS)
    ap.add_argument("--nsets", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--waves-per-run", type=int, default=0)
    ap.add_argument("--runs-per-group", type=int, default=0)
    ap.add_argument("--solve", default="auto", choices=("auto", "reference", "fast"),
                    help="form of the Hines elimination (auto: reciprocal form on well-scaled trees)")
    ap.add_argument("--cpu-sets", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy-peak", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 HBM traffic passes")
    ap.add_argument("--generated", action="store_true",
                    help="simulate with the mechanism set generated from the model's NMODL files "
                         "(ttpc1200: set 3, active500: set 4) instead of the hand-written one")
    ap.add_argument("--copy-peak", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--window", type=int, default=0, help=argparse.SUPPRESS)
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
    if args.pmc_child and args.workload != "mixed4":
        pmc_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (runs are independent, "
              f"so the throughput of N GPUs is N times this figure).", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be positive and --warmup non-negative", file=sys.stderr)
        sys.exit(2)

    from neurogpu_amd import models
    if args.workload == "mixed4":
        return bench_mixed(args)

    w = models.workload(args.workload, nsets=args.nsets, nt=args.nt)
    m = w.model
    nt = int(w.dts.size)
    wins = windows(nt, args.steps)
    wlen = max(b - a for a, b in wins)
    say = lambda msg: print(msg, flush=True)
    say(f"bench: {w.name} N={m.tree.N} runs={w.nruns} nt={nt} S={m.state_count():.2f}; "
        f"{len(wins)} windows of {wlen} steps")
    # measurements in child processes first (rocprofv3 passes, torch copy peak)
    traffic = None
    if not args.no_pmc:
        say("rocprofv3 FETCH_SIZE / WRITE_SIZE passes (child processes)")
        traffic = measured_traffic(args.workload, args.nsets, args.nt, wlen, args.solve, args.generated)
        say(f"traffic per window launch: {traffic['bytes_per_launch']:.4e} B" if traffic else "traffic: unavailable")
    copy_peak = None if args.no_copy_peak else copy_peak_gbs()

    sim = configured_simulator(w, args)
    for i in range(args.warmup):
        a, b = wins[i % len(wins)]
        st = sim.run_steps(a, b)
        say(f"warmup {i}: steps {a}..{b} kernel {st['ms_kernel']:.1f} ms")
    kernel_ms, launch_cs = [], []
    t0 = time.perf_counter()
    for i, (a, b) in enumerate(wins):
        st = sim.run_steps(a, b)
        kernel_ms.append(st["ms_kernel"])
        launch_cs.append(st["comp_steps"])
        say(f"step {i}: steps {a}..{b} kernel {st['ms_kernel']:.1f} ms")
    wall = time.perf_counter() - t0
    gpu = sim.fetch().reshape(w.x.shape[0], w.amps.shape[0], len(m.rec_sites), nt)
    sim.close()

    comp_steps = w.comp_steps
    value = comp_steps / wall
    launch_ms = float(np.mean(kernel_ms))
    cs_launch = float(np.mean(launch_cs))
    bcs = algorithmic_bytes_per_comp_step(m)
    achieved = cs_launch * bcs / (launch_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "kernel": f"{SIM_KERNEL}<{'SetGen_' if args.generated else ''}{m.mechset}>",
                "bytes_per_comp_step": round(bcs, 3),
                "kernel_ms_per_launch": round(launch_ms, 3), "comp_steps_per_launch": cs_launch,
                "launches": len(wins), "steps_per_launch": wlen}
    if traffic:
        roofline["traffic_over_algorithmic"] = round(traffic["bytes_per_launch"] / (cs_launch * bcs), 3)
        roofline["actual_GBs"] = round(traffic["bytes_per_launch"] / (launch_ms * 1e-3) / 1e9, 1)
        roofline["pmc"] = {k: traffic[k] for k in ("FETCH_SIZE", "WRITE_SIZE")}
    if copy_peak:
        roofline["copy_peak_GBs"] = round(copy_peak, 1)

    result = {
        "metric": METRIC, "value": value, "unit": "comp-steps/s", "n_gpus": 1, "steps": len(wins),
        "warmup": args.warmup, "ms_per_step": wall / len(wins) * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.name}: {m.tree.N} compartments, "
                               f"{m.mechset} mechanisms{' generated from their NMODL files' if args.generated else ''}, "
                               f"{w.x.shape[0]} sets x {w.amps.shape[0]} stimuli, dt {w.dts[0]} ms, "
                               f"{nt} steps; a bench step advances every run by one window of about "
                               f"{wlen} dt steps, the timed steps cover the protocol once",
                   "runs": w.nruns, "compartments": m.tree.N, "nt": nt,
                   "launch": {k: st[k] for k in ("waves_per_run", "runs_per_group", "slots_per_lane",
                                                  "runs_per_block")},
                   "solve": {1: "reference order", 2: "reciprocal form"}[st["solve"]]},
        "roofline": roofline,
    }
    if not args.no_cpu:
        threads = cpu_threads()
        cpu_sets = args.cpu_sets if args.cpu_sets is not None else (32 if args.workload == "ttpc1200" else 16)
        nsets = min(cpu_sets, w.x.shape[0])
        say(f"cpu baseline: {nsets} sets on {threads} threads")
        ref, cps, el = cpu_baseline(w, nsets, threads)
        d = float(np.max(np.abs(gpu[:nsets] - ref)))
        same_spikes = bool(np.array_equal(spike_counts(gpu[:nsets]), spike_counts(ref)))
        _, cps1, el1 = cpu_baseline(w, 1, 1)
        result["cpu_baseline"] = {"value": cps, "unit": "comp-steps/s", "cores": threads, "kind": "port",
                                  "sample": f"{nsets} sets x {w.amps.shape[0]} stimuli x {nt} steps "
                                            f"of the same workload ({el:.1f} s, OpenMP over runs)",
                                  "cpu": cpu_model(), "nproc": os.cpu_count(),
                                  "one_thread": {"value": cps1, "sample": f"1 set x {w.amps.shape[0]} stimuli "
                                                                          f"({el1:.1f} s)"}}
        result["parity_sample_max_abs_mV"] = d
        result["parity_sample_spike_counts_equal"] = same_spikes
        if d > 1e-6 or not same_spikes:
            print(f"PARITY FAILURE on the sample: max |dv| {d} mV, spike counts equal: {same_spikes}",
                  file=sys.stderr)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
