"""The drop-in boundary: the reference's Python entry points and the
``neuroGPU [i]`` executable contract (GUI/run_neuroGPU.py:15-48,
GUI/NeuroGPUFromPkl.py:13-19, CANON/kernel.cu:8-34, CANON/Util.cu:443-456)."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from conftest import TOL_MV, TTPC_DIR, spike_counts
from neurogpu_amd import NeuroGPUFromPkl, formats, models, run_neuroGPU
from neurogpu_amd._lib import LIB_PATH

CLI = os.path.join(os.path.dirname(LIB_PATH), "neuroGPU")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _expected_sweep(base):
    rows = []
    for p1 in np.linspace(0, 10, 100):
        for p2 in np.linspace(0, 20, 100):
            r = np.array(base, dtype=float)
            r[0], r[2] = p1, p2
            rows.append(r)
    return np.array(rows)


def test_gen_params_sweep(tmp_path):
    table = tmp_path / "opt_table.csv"
    table.write_text("gna,gk,gca,gl\n0.1,0.2,0.3,0.4\n10,20,30,40\n1.5,2.5,3.5,4.5\n")
    out = tmp_path / "params.csv"
    P = run_neuroGPU.gen_params(str(table), str(out))
    np.testing.assert_array_equal(P, _expected_sweep([1.5, 2.5, 3.5, 4.5]))
    np.testing.assert_allclose(np.loadtxt(out, delimiter=","), P, rtol=1e-15)
    assert P.shape == (10000, 4)


def test_my_gen_params_sweep(tmp_path):
    base = tmp_path / "params.csv"
    base.write_text("0.5 1.5 2.5 3.5 4.5\n")
    P = run_neuroGPU.my_gen_params(str(base), str(tmp_path / "out.csv"))
    np.testing.assert_array_equal(P, _expected_sweep([0.5, 1.5, 2.5, 3.5, 4.5]))


def test_run_params_with_pkl_writes_the_parameter_matrix(tmp_path):
    w = models.workload("active500", nsets=5, nt=2)
    m = w.model
    desc = tmp_path / "model.json"
    data = tmp_path / "Data"
    data.mkdir()
    NeuroGPUFromPkl.save_model_description(m, str(data), str(desc))
    pfile = tmp_path / "new_params.csv"
    np.savetxt(pfile, w.x, delimiter=",")
    P = NeuroGPUFromPkl.run_params_with_pkl(str(desc), str(pfile), len(m.free))
    np.testing.assert_allclose(P, m.allparams(np.loadtxt(pfile, delimiter=",")), rtol=0)
    back = formats.read_allparams(str(data / "AllParams.csv"), m.ncomps, m.nparams)
    np.testing.assert_allclose(back, P, rtol=1e-5)          # "%.5e" in the file
    assert json.load(open(desc))["free"] == [fp.name for fp in m.free]


def test_cli_without_a_device_is_an_error(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([CLI, "0"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "neuroGPU:" in r.stderr


def _data_copy(tmp_path, suffix):
    root = tmp_path / "model"
    data = root / "Data"
    shutil.copytree(TTPC_DIR, data)
    shutil.copy(data / "Stim_raw.csv", data / f"Stim_raw{suffix}.csv")
    (root / "bin").mkdir()
    return root, data


def _oracle_of(data_dir):
    from oracle import cpu_ref
    d = str(data_dir)
    t = formats.read_tree(os.path.join(d, "BasicConstSegP.csv"))
    P = formats.read_allparams(os.path.join(d, "AllParams.csv"), t.ncomps, 19)
    sd = formats.read_sim(os.path.join(d, "Sim.csv"))
    st = formats.read_stim(d)
    return cpu_ref.simulate(t, 1, P, st.amps, st.dts, st.stim_seg, st.area, sd.v_init, sd.rec_sites)


@pytest.mark.gpu
def test_cli_runs_the_reference_data_directory(tmp_path):
    """neuroGPU 0 from bin/ next to Data/: VHotP0.dat and RunTimes.csv as the
    reference writes them, traces within the bar of the oracle."""
    root, data = _data_copy(tmp_path, "0")
    _say("case neuroGPU 0 on the reference TTPC data directory")
    r = subprocess.run([CLI, "0"], cwd=root / "bin", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = _oracle_of(data)
    gpu = formats.read_vhot(str(data / "VHotP0.dat")).reshape(ref.shape)
    d = float(np.max(np.abs(gpu - ref)))
    _say(f"neuroGPU 0: max|gpu-oracle| = {d:.3e} mV")
    assert d <= TOL_MV
    np.testing.assert_array_equal(spike_counts(gpu), spike_counts(ref))
    nsets, ms = open(data / "RunTimes.csv").read().strip().split(",")
    assert int(nsets) == ref.shape[0] and float(ms) > 0


@pytest.mark.gpu
def test_cli_reports_a_missing_file(tmp_path):
    root, data = _data_copy(tmp_path, "0")
    os.remove(data / "Stim_meta.csv")
    r = subprocess.run([CLI, "0"], cwd=root / "bin", capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "Stim_meta" in r.stderr
    assert not (data / "VHotP0.dat").exists()


@pytest.mark.gpu
def test_run_neurogpu_run_matches_the_oracle(tmp_path):
    root, data = _data_copy(tmp_path, "3")
    _say("case run_neuroGPU.run on the reference TTPC data directory")
    out = run_neuroGPU.run(str(data), suffix="3")
    ref = _oracle_of(data)
    assert float(np.max(np.abs(out.reshape(ref.shape) - ref))) <= TOL_MV
    np.testing.assert_array_equal(formats.read_vhot(str(data / "VHotP3.dat")), out)
