"""Parameter refresh without re-porting (the reference's GUI/NeuroGPUFromPkl.py).

The reference reloads pickled porting metadata and re-runs NEURON to write
AllParams.csv (GUI/NeuroGPUFromPkl.py:13-19).  Here the model description is
JSON (no pickle, no NEURON): a template parameter row in which free parameter
i sits at every position holding the sentinel 12300 + i, the convention of
Figures/Figure4_DEAP/extractModel_mappings.py:258-270, plus the data
directory.  ``run_params_with_pkl`` fills the template from a parameter file
and writes AllParams.csv there.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import formats
from .models import SENTINEL


def save_model_description(model, data_dir: str, path: str) -> None:
    """Template row with sentinels at each free parameter's positions."""
    tpl = model.param_template()
    with open(path, "w") as fh:
        json.dump({"template": tpl.tolist(), "ncomps": model.ncomps, "nparams": model.nparams,
                   "free": [fp.name for fp in model.free], "data_dir": data_dir}, fh)


def params_to_allparams(template: np.ndarray, psize: int):
    """Positions of each free parameter (create_params_to_all_params,
    extractModel_mappings.py:258-270)."""
    return {i: np.nonzero(template == SENTINEL + i)[0] for i in range(psize)}


def allparams_from_mapping(template: np.ndarray, params: np.ndarray) -> np.ndarray:
    """Rows of the parameter matrix (my_assemble_allparams, extractModel_mappings.py:158-171)."""
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    idx = params_to_allparams(np.asarray(template), params.shape[1])
    out = np.repeat(np.asarray(template, dtype=np.float64)[None, :], params.shape[0], axis=0)
    for i, cols in idx.items():
        out[:, cols] = params[:, i:i + 1]
    return out


def run_params_with_pkl(model_json: str, new_params_fn: str, psize: int) -> np.ndarray:
    """Fill the template from ``new_params_fn`` (one row per set, comma- or
    space-separated) and write <data_dir>/AllParams.csv; returns the matrix."""
    with open(model_json) as fh:
        desc = json.load(fh)
    raw = np.genfromtxt(new_params_fn, delimiter="," if "," in open(new_params_fn).readline() else None)
    params = np.atleast_2d(raw)[:, :psize] if raw.ndim > 1 or psize > 1 else np.atleast_2d(raw).T
    P = allparams_from_mapping(np.array(desc["template"]), params)
    formats.write_allparams(os.path.join(desc["data_dir"], "AllParams.csv"), P)
    return P
