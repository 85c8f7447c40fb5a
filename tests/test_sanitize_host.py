"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU:
libneurogpu's file parsers, tree checks and level-schedule builder (through
ngpu_open, on the reference fixtures and on malformed copies), the
mechanism-pass schedule builders of schedule.h, the tree-map scheduler of
treemap.h, and the FP64 oracle
(oracle/cpu_ref.c).  Sanitizers instrument host code only; no GPU code is
built with them."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import GOLDEN, ROOT, TTPC_DIR

CSRC = os.path.join(ROOT, "neurogpu_amd", "csrc")
BUILD = os.path.join(ROOT, "neurogpu_amd", "_build")
HIPCC = "/opt/rocm/bin/hipcc"
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def _clean(out):
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]


@pytest.fixture(scope="module")
def open_exe(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("san")
    host = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
            "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
    flags = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-ffp-contract=off"]
    obj = str(d / "ngpu_host_san.o")
    subprocess.run([HIPCC] + flags + host + ["-x", "hip", "-c", os.path.join(CSRC, "ngpu_host.cpp"), "-o", obj],
                   check=True, capture_output=True)
    drv = str(d / "drv.o")
    subprocess.run([HIPCC] + flags + host + ["-c", os.path.join(ROOT, "tests", "sanitize_open.cpp"), "-o", drv],
                   check=True, capture_output=True)
    others = [os.path.join(BUILD, f"{n}.o") for n in ("sim_passive", "sim_ttpc", "sim_mainen", "sim_ttpcgen", "sim_mainengen", "features", "mlp",
                                                        "params")]
    exe = str(d / "sanitize_open")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-fsanitize=address", "-fsanitize=undefined", "-fno-gpu-sanitize",
                    drv, obj] + others + ["-o", exe], check=True, capture_output=True)
    return exe


def _malformed(tmp_path):
    """Copies of the reference TTPC directory, each broken in one way."""
    out = {}

    def copy(name, edit):
        d = tmp_path / name
        shutil.copytree(TTPC_DIR, d)
        edit(d)
        out[name] = str(d)

    def lines(path, fn):
        L = open(path).read().split("\n")
        open(path, "w").write("\n".join(fn(L)))

    copy("blank_meta", lambda d: lines(d / "Stim_meta.csv", lambda L: [L[0], ""] + L[2:]))
    copy("neg_nstim", lambda d: lines(d / "Stim_meta.csv", lambda L: ["-5"] + L[1:]))
    copy("trunc_tree", lambda d: lines(d / "BasicConstSegP.csv", lambda L: L[:9]))
    copy("neg_models", lambda d: lines(d / "BasicConstSegP.csv", lambda L: L[:7] + ["-2,"] + L[8:]))
    copy("short_fidx", lambda d: lines(d / "BasicConstSegP.csv", lambda L: L[:20] + [",".join(L[20].split(",")[:7])] + L[21:]))
    copy("bad_level", lambda d: lines(d / "BasicConstSegP.csv", lambda L: L[:23] + [",".join(["9999"] * len(L[23].split(",")))] + L[24:]))
    copy("bad_ks", lambda d: lines(d / "BasicConstSegP.csv", lambda L: L[:4] + [",".join(["0"] * len(L[4].split(",")))] + L[5:]))
    copy("short_row", lambda d: lines(d / "AllParams.csv", lambda L: [L[0], "1,2,3"] + L[2:]))
    copy("no_times", lambda d: os.remove(d / "times.csv"))
    return out


def test_file_boundary_under_sanitizers(open_exe, tmp_path):
    good = [TTPC_DIR]
    bad = _malformed(tmp_path)
    r = subprocess.run([open_exe] + good + list(bad.values()), capture_output=True, text=True, timeout=300, env=SAN_ENV)
    _clean(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr[-3000:]
    rc = {ln.split("\t")[0]: int(ln.split("\t")[1]) for ln in r.stdout.strip().split("\n")}
    import torch
    if not torch.cuda.is_available():
        assert rc[TTPC_DIR] == -3                     # parsed and checked; no device here
    for name, d in bad.items():
        assert rc[d] in (-1, -2), (name, rc[d])       # a file or argument error, never a crash


def test_schedule_builders_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_schedule")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-I", CSRC, os.path.join(ROOT, "tests", "sanitize_schedule.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=SAN_ENV)
    _clean(r.stdout + r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]


def test_tree_map_scheduler_under_sanitizers(tmp_path):
    """treemap.h (dispersal, maps, records) over 400 random morphologies."""
    exe = str(tmp_path / "sanitize_treemap")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-I", CSRC, os.path.join(ROOT, "tests", "sanitize_treemap.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, str(tmp_path / "maps.csv")], capture_output=True, text=True, timeout=300, env=SAN_ENV)
    _clean(r.stdout + r.stderr)
    assert r.returncode == 0 and "0 bad" in r.stdout, r.stdout + r.stderr[-3000:]


def test_oracle_under_sanitizers(tmp_path):
    """cpu_ref.c built with ASan/UBSan runs small workloads of every
    mechanism set and the reference fixture, and gives the same traces as
    the normal build."""
    lib = str(tmp_path / "libcpuref_san.so")
    subprocess.run(["gcc", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-mfma", "-mavx2",
                    "-fopenmp", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-I", CSRC, os.path.join(ROOT, "oracle", "cpu_ref.c"), "-o", lib, "-lm"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan = subprocess.run(["gcc", "-print-file-name=libubsan.so"], capture_output=True, text=True).stdout.strip()
    script = tmp_path / "run.py"
    script.write_text(f"""
import sys, numpy as np
sys.path.insert(0, {ROOT!r})
from neurogpu_amd import models
from oracle import cpu_ref
out = {{}}
for name, ns, nt in (("passive64", 4, 200), ("active500", 2, 300), ("ttpc1200", 1, 200)):
    w = models.workload(name, nsets=ns, nt=nt)
    m = w.model
    out[name] = cpu_ref.simulate(m.tree, m.mechset_id, w.params, w.amps, w.dts, m.stim_seg, m.stim_area,
                                 m.v_init, m.rec_sites, nthreads=2)
np.savez({str(tmp_path / "san.npz")!r}, **out)
""")
    env = dict(SAN_ENV, LD_PRELOAD=f"{asan}:{ubsan}", NGPU_CPUREF_LIB=lib)
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=600, env=env)
    _clean(r.stdout + r.stderr)
    assert r.returncode == 0, r.stderr[-3000:]
    import numpy as np
    from neurogpu_amd import models
    from oracle import cpu_ref
    san = np.load(tmp_path / "san.npz")
    for name, ns, nt in (("passive64", 4, 200), ("active500", 2, 300), ("ttpc1200", 1, 200)):
        w = models.workload(name, nsets=ns, nt=nt)
        m = w.model
        ref = cpu_ref.simulate(m.tree, m.mechset_id, w.params, w.amps, w.dts, m.stim_seg, m.stim_area, m.v_init,
                               m.rec_sites, nthreads=2)
        np.testing.assert_array_equal(san[name], ref)
