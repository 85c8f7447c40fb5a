"""Python entry points of the reference's GUI/run_neuroGPU.py, on libneurogpu.

``gen_params`` / ``my_gen_params`` build the same 100 x 100 parameter sweep
(parameter 0 over linspace(0, 10, 100), parameter 2 over linspace(0, 20, 100),
the rest at the base row) as GUI/run_neuroGPU.py:15-48; ``run`` replaces
spawning ``neuroGPU`` on a data directory; ``write_data_dir`` writes every
input file of the reference's Data/ layout for a workload.
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import formats, tree
from ._lib import Simulator


def _sweep(base: np.ndarray) -> np.ndarray:
    rows = []
    for p1 in np.linspace(0, 10, 100):
        for p2 in np.linspace(0, 20, 100):
            cur = np.array(base, dtype=np.float64, copy=True)
            cur[0] = p1
            cur[2] = p2
            rows.append(cur)
    return np.array(rows)


def gen_params(opt_table_fn: str, out_fn: str) -> np.ndarray:
    """opt_table.csv: header of names, rows lower bound, upper bound, base
    (GUI/run_neuroGPU.py:15-30).  Writes and returns the sweep."""
    table = np.genfromtxt(opt_table_fn, dtype=float, delimiter=",", names=True)
    base = np.array(table[2].tolist(), dtype=np.float64)
    params = _sweep(base)
    np.savetxt(out_fn, params, delimiter=",")
    return params


def my_gen_params(base_fn: str, out_fn: str) -> np.ndarray:
    """Same sweep around a whitespace-separated base row (GUI/run_neuroGPU.py:32-48)."""
    params = _sweep(np.genfromtxt(base_fn))
    np.savetxt(out_fn, params, delimiter=",")
    return params


def run(data_dir: str, suffix: str = "", device: int = 0, write: bool = True) -> np.ndarray:
    """Simulate a reference data directory on the GPU; writes VHotP<suffix or 0>.dat
    like ``neuroGPU <suffix>`` and returns the traces as a flat array."""
    sim = Simulator.open(data_dir, suffix=suffix, device=device)
    try:
        out, _ = sim.run()
        if write:
            sim.write_vhot(os.path.join(data_dir, f"VHotP{suffix or 0}.dat"))
    finally:
        sim.close()
    return out


def write_data_dir(workload, data_dir: str, suffix: str = "", rec_sites: Optional[np.ndarray] = None) -> None:
    """All input files of the reference's Data/ directory for a workload."""
    os.makedirs(data_dir, exist_ok=True)
    m = workload.model
    tree.write_tree(os.path.join(data_dir, "BasicConstSegP.csv"), m.tree)
    formats.write_allparams(os.path.join(data_dir, "AllParams.csv"), workload.params)
    rec = m.rec_sites if rec_sites is None else rec_sites
    formats.write_sim(os.path.join(data_dir, "Sim.csv"), formats.SimData(m.v_init, float(workload.dts[0]), rec))
    st = formats.StimData(workload.amps.shape[0], workload.dts.size, 0, m.stim_seg + 1, m.stim_area,
                          workload.amps, workload.dts)
    formats.write_stim(data_dir, st, suffix)
