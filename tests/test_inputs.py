"""Input layer: the in.ether tags, the four input files, and the failure
conditions the reference terminates on (src/read_input.f90:57-146,
src/j_values.f90:31-37, src/get_sia_values.f90:64-78, src/read_structure.f90:103-106,
src/generate_supercell.f90:99-101), for the Python reader and the Fortran driver."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ether_amd import workloads
from ether_amd.inputs import EtherInputError, parse_in_ether, read_directory

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "tests", "golden", "examples")
DRIVER = os.path.join(ROOT, "ether_amd", "lib", "ether_gpu_driver")


@pytest.mark.parametrize("name", ["Pyrochlore", "Triangular", "SimpleCubic_FM", "SimpleCubic_AFM",
                                  "SimpleCubic_333", "SimpleCubic_141414"])
def test_examples_parse(name):
    m = read_directory(os.path.join(EX, name))
    assert m.n_sites == int(np.prod(m.sc)) * len(m.basis_species)
    assert m.tags.tmcs > m.tags.tmcs_eq


def test_template_with_unknown_tag_terminates():
    # inputfiles/in.ether carries `dope`, which read_input rejects
    with pytest.raises(EtherInputError, match="unknown"):
        read_directory(os.path.join(EX, "inputfiles"))


def test_defaults_and_tag_values():
    t = parse_in_ether("model = XYZ\nangle = .T. 50 ! comment\n#x\n\tcoa = 5\n", 1)
    assert t.model == "xyz" and t.angle and abs(t.dphi - np.radians(50)) < 1e-15
    assert t.coa == 5 and t.tmcs == 5000 and t.tmcs_eq == 3000 and t.seed == 1992
    assert t.sc == (2, 2, 2) and t.nbdfc == 1e-5 and t.exchange_interval == 11
    assert len(t.temperatures()) == 19


@pytest.mark.parametrize("text,match", [
    ("mcs = 100 200\n", "EQUILIBRATIONS"),
    ("model ising\n", "Missing '='"),
    ("para = .T. 0\n", "cannot be zero"),
    ("bogus = 1\n", "unknown"),
])
def test_tag_errors(text, match):
    with pytest.raises(EtherInputError, match=match):
        parse_in_ether(text, 1)


def test_workload_temperature_ladders():
    assert len(workloads.model("tri_afm_96", sc="3 3 1").tags.temperatures()) == 32
    T = workloads.model("pyro_pt_16", sc="1 1 1").tags.temperatures()
    assert len(T) == 64 and T[0] == 1.0 and abs(T[-1] - 0.05) < 1e-12


def test_units_and_scalings():
    m = workloads.model("tri_afm_96", sc="3 3 1", spin="2")
    # J in meV -> eV, scaled by s_a s_b; D by s^2 (src/j_values.f90:83, get_sia_values.f90:58)
    assert m.j_exc[0, 0, 0, 0] == pytest.approx(4.0 / 1000)
    assert m.sia_vec[0, 2] == pytest.approx(-0.1 * 4 / 1000)
    assert m.mu_b == pytest.approx(5.7883818060e-5) and m.g_factor == 2.0
    p = workloads.model("tri_afm_96", sc="3 3 1", para=".T. 2")
    assert p.j_exc[0, 0, 0, 0] == pytest.approx(0.5) and p.kb == 1.0 and p.g_factor == 1.0
    assert p.h[2] == pytest.approx(0.5)


def _driver(tmp_path, files):
    for k, v in files.items():
        (tmp_path / k).write_text(v)
    return subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
@pytest.mark.parametrize("mutate,match", [
    (lambda f: f.update({"in.ether": f["in.ether"] + "dope = .F.\n"}), "unknown/missing"),
    (lambda f: f.pop("j_exchange"), "j_exchange"),
    (lambda f: f.pop("in.ether"), "in.ether"),
    (lambda f: f.update({"in.ether": f["in.ether"].replace("mcs = 5000 3000", "mcs = 10 20")}),
     "EQUILIBRATIONS"),
    (lambda f: f.pop("single_ion_anisotropy"), "single_ion_anisotropy"),
    (lambda f: f.update({"in.ether": f["in.ether"] + "species = Zz\n"}), "not present"),
])
def test_fortran_driver_terminates_like_the_reference(tmp_path, mutate, match):
    f = workloads.files("tri_afm_96", sc="3 3 1")
    f = {{"structure": "structure.vasp"}.get(k, k): v for k, v in f.items()}
    mutate(f)
    r = _driver(tmp_path, f)
    assert r.returncode != 0
    assert match in r.stdout + r.stderr
