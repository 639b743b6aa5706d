"""gpu.ether settings, the learned ladder in both drivers, and batch manifests.

Launch shape, lanes per site, graph replay and chunking change no result, so
a run directory with such a gpu.ether must give bit-identical moments to the
same directory without it, in both drivers.  A batch manifest's runs must
write the same files as each run directory run alone.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from ether_amd import driver, outputs, workloads
from ether_amd.inputs import read_directory
from tests.helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ether_amd", "lib", "ether_gpu_driver")


def _obs(path, m):
    return outputs.read_direct(os.path.join(path, "ETHER.obs"), outputs.TOTAL_OBSERVABLES + 3 * m.n_species)


def test_engine_settings_change_no_result(tmp_path):
    over = dict(sc="6 6 6", mcs="300 150", repeat="3", temp="60 6 18")
    a, b = tmp_path / "a", tmp_path / "b"
    workloads.write_dir("sc_heis_afm_32", str(a), **over)
    workloads.write_dir("sc_heis_afm_32", str(b), **over)
    (b / "gpu.ether").write_text("launch = 2 0\nlanes = 64\ngraphs = .F.\nchunk = 37\n")
    say("gpu.ether settings: python driver")
    ra, rb = driver.run(read_directory(str(a))), driver.run(read_directory(str(b)))
    assert np.array_equal(ra.mean, rb.mean) and np.array_equal(ra.sem, rb.sem)
    assert np.array_equal(ra.final_S, rb.final_S)
    rb.write(str(b / "py"))
    say("gpu.ether settings: fortran driver")
    r = subprocess.run([DRIVER], cwd=b, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = ra.model
    assert np.array_equal(np.fromfile(b / "ETHER.spn", "<f8"), np.fromfile(b / "py" / "ETHER.spn", "<f8"))
    assert np.allclose(_obs(str(b), m), _obs(str(b / "py"), m), rtol=1e-12, atol=1e-14)


def test_learned_ladder_through_the_driver(tmp_path):
    # 128-spin pyrochlore, 16 temperatures on 20 -> 1 K, training in the equilibration
    workloads.write_dir("pyro_pt_16", str(tmp_path), sc="2 2 2", mcs="6000 3000", repeat="2",
                        temp="20 1 1.2666666666666666", pt=".T. 3")
    (tmp_path / "gpu.ether").write_text("ladder = learned 2000 4\n")
    say("learned ladder via python -m ether_amd")
    r = subprocess.run([sys.executable, "-m", "ether_amd", str(tmp_path)], cwd=ROOT,
                       capture_output=True, text=True, timeout=600)
    say(r.stdout[-300:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ladder " in r.stdout
    rows = [ln for ln in open(tmp_path / "energy.dat") if not ln.lstrip().startswith("#")]
    T = np.array([float(ln.split()[0]) for ln in rows])
    assert len(T) == 16 and np.all(np.diff(T) < 0)
    assert T[0] == pytest.approx(20.0) and T[-1] == pytest.approx(1.0)
    acc = np.array([float(ln.split()[-1]) for ln in rows])
    assert np.all((acc >= 0) & (acc <= 100)) and acc[0] > acc[-1]
    # the ETHER.obs temperatures are the ladder that ran production
    m = read_directory(str(tmp_path))
    assert np.allclose(_obs(str(tmp_path), m)[:, 0], T, atol=5e-6)
    # the Fortran driver runs the same training (ether_ladder_train) and
    # validation, so it writes the same ladder and numbers
    fdir = tmp_path / "fortran"
    workloads.write_dir("pyro_pt_16", str(fdir), sc="2 2 2", mcs="6000 3000", repeat="2",
                        temp="20 1 1.2666666666666666", pt=".T. 3")
    (fdir / "gpu.ether").write_text("ladder = learned 2000 4\n")
    say("learned ladder via the Fortran driver")
    r = subprocess.run([DRIVER], cwd=fdir, capture_output=True, text=True, timeout=600)
    say(r.stdout[-300:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "learned tempering ladder" in r.stdout
    for f in ("energy.dat", "magnetization.dat"):
        a = np.array([[float(x) for x in ln.split()] for ln in open(tmp_path / f) if not ln.lstrip().startswith("#")])
        b = np.array([[float(x) for x in ln.split()] for ln in open(fdir / f) if not ln.lstrip().startswith("#")])
        assert a.shape == b.shape and np.allclose(a, b, rtol=2e-5, atol=2e-5), f


def test_batch_manifest_equals_runs_alone(tmp_path):
    runs = {
        "ising": ("sc_ising_16", dict(sc="6 6 6", mcs="300 150", repeat="3", temp="100 10 30")),
        "heis": ("sc_heis_afm_32", dict(sc="4 4 4", mcs="300 150", repeat="2", temp="60 6 27",
                                        ssp=".T.", ss_dis="1.01", coa="10", ovrr=".F. 0 0")),
        "heis2": ("sc_heis_afm_32", dict(sc="6 4 4", mcs="300 150", repeat="2", temp="30 6 12",
                                         seed="7", coa="10", ovrr=".F. 0 0")),
        "pyro": ("pyro_pt_16", dict(sc="2 2 2", mcs="330 110", repeat="2", temp="1 0.05 0.19")),
    }
    for d, (name, over) in runs.items():
        workloads.write_dir(name, str(tmp_path / "batch" / d), **over)
        workloads.write_dir(name, str(tmp_path / "alone" / d), **over)
    (tmp_path / "batch.ether").write_text("".join(f"batch/{d}\n" for d in runs))
    say("batch manifest")
    r = subprocess.run([sys.executable, "-m", "ether_amd", "--batch", str(tmp_path / "batch.ether")],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    say(r.stdout[-400:])
    assert r.returncode == 0, r.stdout + r.stderr
    for d in runs:
        m = read_directory(str(tmp_path / "alone" / d))
        res = driver.run(m)
        res.write(str(tmp_path / "alone" / d))
        a, b = tmp_path / "alone" / d, tmp_path / "batch" / d
        assert np.array_equal(np.fromfile(a / "ETHER.spn", "<f8"), np.fromfile(b / "ETHER.spn", "<f8")), d
        oa, ob = _obs(str(a), m), _obs(str(b), m)
        assert np.array_equal(oa, ob), d
        for f in ("energy.dat", "magnetization.dat", "spiral_state_parameters.dat", "gss.dat"):
            assert open(a / f).read() == open(b / f).read(), (d, f)
