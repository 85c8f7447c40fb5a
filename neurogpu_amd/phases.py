"""Per-phase cycles of the simulation kernel, from the profiling build.

    make -C neurogpu_amd/csrc phases
    python -m neurogpu_amd.phases [--workload ttpc1200] [--nsets 2048] [--nt 600] [--solve auto]

The profiling build is libneurogpu with -DNG_PHASE_CYCLES=1
(neurogpu_amd/_build_phases/): thread 0 of every workgroup adds the cycles of
each phase of every step.  Prints, per step of one workgroup: the mechanism
pass, the elimination (BeforeLU), the pivot quotients, the
back-substitution and the v/dv update with the trace, as cycles and shares.
"""
import argparse
import ctypes as C
import json
import os
import sys

PHASES = ("mechanism pass", "elimination", "pivot quotients", "back-substitution", "v/dv update and trace")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ttpc1200")
    ap.add_argument("--nsets", type=int, default=2048)
    ap.add_argument("--nt", type=int, default=600)
    ap.add_argument("--solve", default="auto", choices=("auto", "reference", "fast"))
    args = ap.parse_args(argv)
    here = os.path.dirname(os.path.abspath(__file__))
    lib = os.path.join(here, "_build_phases", "libneurogpu.so")
    if not os.path.exists(lib):
        print(f"{lib} is missing: make -C neurogpu_amd/csrc phases", file=sys.stderr)
        return 2
    os.environ["NGPU_LIB"] = lib
    from neurogpu_amd import _lib, models
    L = _lib.lib()
    w = models.workload(args.workload, nsets=args.nsets, nt=args.nt)
    fn = getattr(L, f"ngpu_phase_cycles_{w.model.mechset}")
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(C.c_ulonglong), C.c_int]
    buf = (C.c_ulonglong * 8)()
    sim = _lib.workload_simulator(w)
    sim.set_solve({"auto": 0, "reference": 1, "fast": 2}[args.solve])
    sim.run(fetch=False)                         # warm-up
    if fn(buf, 1):
        raise RuntimeError("reading the phase counters failed")
    _, st = sim.run(fetch=False)
    if fn(buf, 1):
        raise RuntimeError("reading the phase counters failed")
    sim.close()
    steps = float(buf[5])
    per = [buf[k] / steps for k in range(5)]
    tot = sum(per)
    out = {"workload": f"{w.name}: {w.nruns} runs x {args.nt} steps", "solve": args.solve,
           "launch": {k: st[k] for k in ("waves_per_run", "runs_per_group", "runs_per_block")},
           "kernel_ms": st["ms_kernel"], "comp_steps_per_s": st["comp_steps_per_s"],
           "cycles_per_step_per_workgroup": round(tot),
           "phases": {n: {"cycles": round(c), "share": round(c / tot, 4)} for n, c in zip(PHASES, per)}}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
