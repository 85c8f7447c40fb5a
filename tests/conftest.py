import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
TTPC_DIR = os.path.join(GOLDEN, "ttpc512")
TREE_FILES = {
    "ttpc512": os.path.join(TTPC_DIR, "BasicConstSegP.csv"),
    "passive384": os.path.join(GOLDEN, "trees", "passive384_BasicConstSegP.csv"),
    "mainen384": os.path.join(GOLDEN, "trees", "mainen384_BasicConstSegP.csv"),
    "cc224": os.path.join(GOLDEN, "trees", "cc224_BasicConstSegP.csv"),
}
TOL_MV = 1e-6   # parity bar: max-abs trace difference against the FP64 oracle


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libneurogpu's HIP kernels)")


def spike_counts(traces, thr=-30.0):
    tr = np.asarray(traces)
    return np.sum((tr[..., 1:] >= thr) & (tr[..., :-1] < thr), axis=-1)
