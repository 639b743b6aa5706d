"""End-to-end runs through both drivers of the same C-ABI.

The Fortran driver (ether_amd/lib/ether_gpu_driver, iso_c_binding) and the
Python driver (ether_amd.driver, ctypes) read the same run directory and use
the same seed.  Both take the generate_supercell start state from the
library's ether_start_state, a function of SEED, so for Ising, Marsaglia and
ANGLE runs alike both drivers must write identical energy.dat /
magnetization.dat numbers (and
spiral_state_parameters.dat with the spiral state on), with identical header
records, and identical ETHER.spn / ETHER.obs records, gss.dat,
moment_vectors.dat and spK.xsf files.
A simple-cubic Heisenberg ferromagnet run also checks the physics: the energy
per spin is monotone in T and approaches the ground-state value 3J at low T.
"""
import os
import subprocess

import numpy as np
import pytest

from ether_amd import driver, outputs, workloads
from ether_amd.inputs import read_directory
from tests.helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "ether_amd", "lib", "ether_gpu_driver")


def _table(path):
    return np.array([[float(x) for x in ln.split()] for ln in open(path) if not ln.lstrip().startswith("#")])


@pytest.mark.parametrize("name,over", [
    ("sc_ising_16", dict(sc="6 6 6", mcs="400 200", repeat="3", temp="100 10 30")),
    ("sc_heis_afm_32", dict(sc="6 6 6", mcs="400 200", repeat="3", temp="60 6 18",
                            ssp=".T.", ss_dis="1.01")),
    ("pyro_pt_16", dict(sc="2 2 2", mcs="440 220", repeat="2", temp="1 0.05 0.19")),
    # ANGLE moves: the start phi feeds fresh_spins' angle move
    ("tri_afm_96", dict(sc="12 12 1", mcs="400 200", repeat="2", temp="20 0.5 6.5")),
])
def test_fortran_and_python_drivers_agree(tmp_path, name, over):
    say(f"drivers {name}")
    workloads.write_dir(name, str(tmp_path), **over)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    say(r.stdout[-400:])
    assert r.returncode == 0, r.stdout + r.stderr
    if name == "tri_afm_96":
        # the same SEED twice: the same start state, the same files
        again = tmp_path / "again"
        workloads.write_dir(name, str(again), **over)
        r = subprocess.run([DRIVER], cwd=again, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        for f in ("energy.dat", "magnetization.dat", "ETHER.spn"):
            assert open(tmp_path / f, "rb").read() == open(again / f, "rb").read(), f
    res = driver.run(read_directory(str(tmp_path)))
    res.write(str(tmp_path / "py"))
    files = ["energy.dat", "magnetization.dat"]
    if "ssp" in over:
        files.append("spiral_state_parameters.dat")
    else:
        assert os.path.getsize(tmp_path / "spiral_state_parameters.dat") == 0
    for f in files:
        a = _table(tmp_path / f)
        b = _table(tmp_path / "py" / f)
        assert a.shape == b.shape and a.shape[0] > 0
        assert np.allclose(a, b, rtol=2e-5, atol=2e-5), f
        # header records are the reference's A13 edits of character(len=20) labels
        ha = open(tmp_path / f, encoding="utf-8").readline()
        hb = open(tmp_path / "py" / f, encoding="utf-8").readline()
        assert ha == hb, f
    _same_state_files(tmp_path, res)


@pytest.mark.parametrize("name,over", [
    ("sc_ising_16", dict(sc="6 6 6", mcs="300 150", repeat="3", temp="100 10 30")),
    ("tri_afm_96", dict(sc="12 12 1", mcs="300 150", repeat="3", temp="20 0.5 6.5")),   # ANGLE
    ("pyro_pt_16", dict(sc="2 2 2", mcs="330 110", repeat="2", temp="1 0.05 0.19")),    # tempering
])
def test_chained_repeats_in_both_drivers(tmp_path, name, over):
    """compat = 8 (ETHER_COMPAT_CHAIN_REPEATS): the repeats run one after
    another on nscan replicas, each from fresh_spins of the previous repeat's
    final state (src/ether.f90:152-155).  Both drivers write the same numbers;
    every repeat books its own samples; the run differs from the default one,
    where every repeat starts from the generate_supercell state."""
    say(f"chained repeats {name}")
    f, d = tmp_path / "f", tmp_path / "default"
    workloads.write_dir(name, str(f), **over)
    workloads.write_dir(name, str(d), **over)
    (f / "gpu.ether").write_text("compat = 8\n")
    r = subprocess.run([DRIVER], cwd=f, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = read_directory(str(f))
    res = driver.run(m)
    res.write(str(f / "py"))
    for fn in ("energy.dat", "magnetization.dat"):
        a = _table(f / fn)
        b = _table(f / "py" / fn)
        assert a.shape == b.shape and a.shape[0] > 0
        assert np.allclose(a, b, rtol=2e-5, atol=2e-5), fn
    assert np.array_equal(np.fromfile(f / "ETHER.spn", "<f8"), np.fromfile(f / "py" / "ETHER.spn", "<f8"))
    # every repeat sampled every coa-th production MCS
    n_samples = (m.tags.tmcs - m.tags.tmcs_eq) // int(m.tags.coa)
    per = res.per_repeat
    assert per.shape[0] == m.tags.repeat
    res_d = driver.run(read_directory(str(d)))
    assert not np.array_equal(res.per_repeat, res_d.per_repeat)
    assert n_samples > 0


def _same_state_files(tmp_path, res):
    """Files built from the final configurations: both drivers read the same
    device state, so the records agree bit for bit and the text records byte
    for byte (the Fortran driver formats them with the reference's own format
    statements; list-directed xsf header lines are compared as numbers)."""
    m = res.model
    N, nscan = m.n_sites, len(res.temps)
    a = np.fromfile(tmp_path / "ETHER.spn", "<f8")
    b = np.fromfile(tmp_path / "py" / "ETHER.spn", "<f8")
    assert a.size == nscan * N * 9 and np.array_equal(a, b)
    nrec = outputs.TOTAL_OBSERVABLES + 3 * m.n_species
    a = outputs.read_direct(str(tmp_path / "ETHER.obs"), nrec)
    b = outputs.read_direct(str(tmp_path / "py" / "ETHER.obs"), nrec)
    assert a.shape == (nscan, nrec)
    assert np.array_equal(a[:, 20:], b[:, 20:])                 # moment vectors: same sums
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14)            # estimators: FP64 in two languages
    for f in ("gss.dat", "moment_vectors.dat"):
        ta = open(tmp_path / f, encoding="utf-8").read()
        tb = open(tmp_path / "py" / f, encoding="utf-8").read()
        assert ta == tb, f
    for T in res.temps:
        f = outputs.spk_name(T)
        la = open(tmp_path / f).read().splitlines()
        lb = open(tmp_path / "py" / f).read().splitlines()
        assert len(la) == len(lb) == 11 + N
        assert la[11:] == lb[11:]
        for x, y in zip(la[:11], lb[:11]):
            assert x.split()[0].isalpha() == y.split()[0].isalpha()
            if not x.split()[0].isalpha():
                assert np.allclose([float(v) for v in x.split()], [float(v) for v in y.split()])


def test_heisenberg_ferromagnet_energy_curve():
    say("heisenberg fm curve")
    m = workloads.model("sc_heis_afm_32", sc="8 8 8", mcs="1500 500", repeat="4",
                        temp="42 2 4", para=".F. 1", cb=".F.")
    m.j_exc = -m.j_exc          # ferromagnet
    res = driver.run(m)
    e = res.mean[:, 0]
    Jev = abs(m.j_exc[0, 0, 0, 0])
    assert len(e) == 11 and res.temps[-1] == pytest.approx(2.0)
    assert np.all(np.diff(e) < 1e-4 * Jev)          # colder is lower in energy
    # e = E_tot/(2N) -> 3J at T -> 0 (J = -1 meV); classical spin waves add
    # kB T per spin (two quadratic modes) at low T
    assert e[-1] == pytest.approx(-3 * Jev + m.kb * 2.0, rel=0.02)
    assert np.all(res.accept_pct > 0)
