"""The C-ABI library loads and exports every entry point of include/neurogpu.h
(no GPU needed: nothing here launches work)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from neurogpu_amd import _lib


def _declared():
    hdr = open(os.path.join(ROOT, "include", "neurogpu.h")).read()
    return sorted(set(re.findall(r"\b(ngpu_[a-z_0-9]+)\s*\(", hdr)))


def test_header_declarations_match_binding():
    assert set(_declared()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_mechset_sizes_without_a_device():
    L = _lib.lib()
    assert [L.ngpu_mechset_nparams(i) for i in range(5)] == [2, 19, 30, 19, 30]
    assert L.ngpu_mechset_nparams(7) < 0 and L.ngpu_mechset_nparams(-1) < 0


def test_missing_device_is_a_clean_error():
    """Without a usable device ngpu_create returns an error code and a message
    (it never falls back to a CPU path)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.NgpuError) as ei:
        _lib.Simulator(0)
    assert ei.value.code < 0


def test_no_oracle_in_product_sources():
    """The product path never links or loads the oracle."""
    pkg = os.path.join(ROOT, "neurogpu_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, fn)).read()
                for bad in ("cpu_ref", "libcpuref", "import oracle", "from oracle"):
                    assert bad not in txt, (fn, bad)
