"""Autotuner of the batch launch: per morphology, waves per run (which sets the
runs per 8-wave workgroup), level-sync granularity (per wave or per workgroup)
and LDS staging of the Hines row coefficients and the run's varying parameters.

Each candidate runs a probe batch of the morphology alone: enough runs to
give every CU four items of the largest shape (parameter sets repeated), for
a few hundred steps, and the candidate with the highest compartment-steps/s
wins.
Choices are cached in a JSON file keyed by (tree hash, mechanism set,
architecture), so a morphology is tuned once per machine.

No reference counterpart: the reference launch shape is fixed
(GUI/BBP_TTPC_EXAMPLE/pyNeuroGPU_unix/src/CudaStuff.cu:487-507).
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib, models

ARCH = "gfx950"
BATCH_NSLOT = 6                 # positions per lane of the batch kernel
BATCH_WAVES = 8                 # waves per workgroup of the batch kernel
CUS = 256                       # compute units of an MI355X
KERNEL_REV = 6                  # bump when the kernel changes: cached choices are per revision
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "neurogpu_amd", "batch_tuning.json")


@dataclasses.dataclass(frozen=True)
class Choice:
    waves_per_run: int
    block_sync: int
    lds_staging: int

    def as_dict(self) -> dict:
        return dataclasses.asdict(self)


def candidates(n_segments: int) -> List[Choice]:
    """Every launch shape the batch kernel supports for a tree of this size."""
    out = []
    for w in (1, 2, 4, 8):
        if 64 * w * BATCH_NSLOT < n_segments:
            continue
        for sync in ((0, 1) if w == 1 else (1,)):
            for stage in (1, 0):
                out.append(Choice(w, sync, stage))
    return out


def tree_key(w: models.Workload) -> str:
    t = w.model.tree
    h = hashlib.sha1()
    for a in (t.ks, t.e, t.f, t.cm, t.bool_model, t.seg_to_comp):
        h.update(np.ascontiguousarray(a).tobytes())
    return f"{h.hexdigest()[:16]}:{w.model.mechset}:{ARCH}:{KERNEL_REV}"


def _probe_workload(w: models.Workload, runs: int, steps: int) -> models.Workload:
    """runs runs of w's parameter sets (repeated when w has fewer), so that a
    probe fills every CU several times whatever the member's own size: the
    batch shares the CUs among its members, so a shape is judged by its
    throughput per CU, not by how fast the member alone would finish."""
    nst = w.amps.shape[0]
    nsets = max(1, -(-runs // nst))
    steps = min(steps, w.dts.size)
    x = w.x[np.arange(nsets) % w.x.shape[0]]
    return models.Workload(w.name, w.model, x, w.amps[:, :steps], w.dts[:steps])


def measure(w: models.Workload, choice: Choice, device: int = 0) -> float:
    """Compartment-steps/s of one batch launch of w alone with this choice."""
    sim = _lib.workload_simulator(w, device)
    try:
        with _lib.Batch(device) as b:
            i = b.add(sim)
            b.set_launch(i, **choice.as_dict())
            st = b.run()
    finally:
        sim.close()
    return st["comp_steps"] / (st["ms_kernel"] * 1e-3)


def item_runs(choice: Choice) -> int:
    """Most runs in one item of this shape: 8 / waves groups of up to four
    runs (one-wave groups) or six (larger groups), as the library allows."""
    return BATCH_WAVES // choice.waves_per_run * (4 if choice.waves_per_run == 1 else 6)


def tune(w: models.Workload, device: int = 0, probe_runs: Optional[int] = None, probe_steps: int = 200,
         cache: Optional[str] = DEFAULT_CACHE, log=None) -> Tuple[Choice, Dict[str, float]]:
    """Best launch choice for w (from the cache when present)."""
    key = tree_key(w)
    db = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
    if key in db:
        return Choice(**db[key]["choice"]), db[key]["rates"]
    cands = candidates(w.model.tree.N)
    if probe_runs is None:
        # four items of the largest shape per CU
        probe_runs = 4 * CUS * max([item_runs(c) for c in cands] or [1])
    probe = _probe_workload(w, probe_runs, probe_steps)
    rates: Dict[str, float] = {}
    best, best_rate = None, -1.0
    for c in cands:
        try:
            measure(probe, c, device)            # first launch builds the tables
            r = measure(probe, c, device)
        except _lib.NgpuError as e:              # e.g. LDS too small for this shape
            if log:
                log(f"  {w.name} {c}: not possible ({e})")
            continue
        rates[json.dumps(c.as_dict(), sort_keys=True)] = r
        if log:
            log(f"  {w.name} {c}: {r:.3e} comp-steps/s")
        if r > best_rate:
            best, best_rate = c, r
    if best is None:
        raise RuntimeError(f"no launch shape fits {w.name} (N={w.model.tree.N})")
    if cache:
        db[key] = {"choice": best.as_dict(), "rates": rates, "name": w.name}
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return best, rates
