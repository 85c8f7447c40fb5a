"""ngpu_open on well-formed and malformed data directories, on the CPU.

ngpu_open parses and checks every file before it sets up the device, so a
malformed directory gives NGPU_E_FILE with a message on any machine (no C++
exception crosses the C ABI, no abort), and a well-formed one gets as far as
the device (NGPU_E_HIP here, where there is none).  File layouts: the
reference's Data directory (CANON/MainC.cu:335-551, CANON/Util.cu:337-432)."""
import ctypes as C
import os
import shutil

import pytest

from conftest import TTPC_DIR
from neurogpu_amd import _lib

NGPU_E_FILE, NGPU_E_HIP = -2, -3


def _no_gpu():
    import torch
    return not torch.cuda.is_available()


def _open(d):
    L = _lib.lib()
    h = C.c_void_p()
    rc = L.ngpu_open(str(d).encode(), b"", -1, 0, C.byref(h))
    msg = (L.ngpu_last_error(None) or b"").decode()
    if rc == 0:
        L.ngpu_close(h)
    return rc, msg


@pytest.fixture
def data(tmp_path):
    d = tmp_path / "Data"
    shutil.copytree(TTPC_DIR, d)
    return d


def _edit(path, fn):
    lines = open(path).read().split("\n")
    open(path, "w").write("\n".join(fn(lines)))


@pytest.mark.skipif(not _no_gpu(), reason="needs a machine without a GPU")
def test_wellformed_directory_reaches_the_device(data):
    rc, msg = _open(data)
    assert rc == NGPU_E_HIP and "no HIP device" in msg


@pytest.mark.parametrize("case", [
    "blank_stim_line", "negative_nstim", "zero_nt", "bad_area", "loc_outside", "missing_stim_meta",
    "truncated_tree", "negative_nmodels", "short_e_line", "huge_n", "params_header", "params_row",
    "recsite_outside", "times_short",
])
def test_malformed_directory_is_a_file_error(data, case):
    meta, tree = data / "Stim_meta.csv", data / "BasicConstSegP.csv"
    if case == "blank_stim_line":
        _edit(meta, lambda L: [L[0], ""] + L[2:])
    elif case == "negative_nstim":
        _edit(meta, lambda L: ["-3"] + L[1:])
    elif case == "zero_nt":
        _edit(meta, lambda L: [L[0], "0"] + L[2:])
    elif case == "bad_area":
        _edit(meta, lambda L: L[:4] + ["nan"] + L[5:])
    elif case == "loc_outside":
        _edit(meta, lambda L: L[:3] + ["100000"] + L[4:])
    elif case == "missing_stim_meta":
        os.remove(meta)
    elif case == "truncated_tree":
        _edit(tree, lambda L: L[:12])
    elif case == "negative_nmodels":
        _edit(tree, lambda L: L[:7] + ["-13,"] + L[8:])
    elif case == "short_e_line":
        _edit(tree, lambda L: L[:2] + [",".join(L[2].split(",")[:10])] + L[3:])
    elif case == "huge_n":
        _edit(tree, lambda L: ["40000000,"] + L[1:])
    elif case == "params_header":
        _edit(data / "AllParams.csv", lambda L: ["-1,"] + L[1:])
    elif case == "params_row":
        _edit(data / "AllParams.csv", lambda L: [L[0], ",".join(L[1].split(",")[:5])] + L[2:])
    elif case == "recsite_outside":
        _edit(data / "Sim.csv", lambda L: L[:3] + ["9999,"] + L[4:])
    elif case == "times_short":
        _edit(data / "times.csv", lambda L: [",".join(L[0].split(",")[:100])] + L[1:])
    rc, msg = _open(data)
    assert rc == NGPU_E_FILE, (rc, msg)
    assert msg


def test_missing_directory_is_a_file_error(tmp_path):
    rc, msg = _open(tmp_path / "nowhere")
    assert rc == NGPU_E_FILE and "cannot open" in msg
