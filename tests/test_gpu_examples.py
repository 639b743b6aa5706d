"""The reference's own example directories, run unchanged.

examples/SimpleCubic/{FM,AFM} (kept under tests/golden/examples): two
species A and B on alternating sites of a simple-cubic lattice with 1 A
spacing (4 + 4 sites in a 2 A cell, 4x4x4 cells), |J| = 3 meV between
unlike neighbours only, ANGLE moves (50 degrees), OVRR 0.6 every 12 MCS,
40 temperatures 80 -> 2 K x 10 repeats, MCS 8000/4000.  The AFM directory
adds stg = .T. with the `staggered` file (-1 on B).  `python -m ether_amd`
writes every result file; at 2 K (kT/|J| = 0.06) the physics fixes:
  - |M| per site -> 1 (FM), staggered |M| -> 1 (AFM), minus the classical
    spin-wave reduction of order kT/(z|J|) and what a quench from a random
    state leaves in 4000 MCS;
  - e = E_tot/(2N) -> -z|J|/2 = -9 meV per site, plus kT per spin for the two
    quadratic modes of each classical spin;
  - per species, moment_vectors.dat gives |M_sp|/N_sp -> 1, and the two
    species' vectors are parallel (FM) or, with the stagger, parallel too (AFM).
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from ether_amd import outputs
from ether_amd.inputs import KB_EV, read_directory
from tests.helpers import say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "tests", "golden", "examples")


def _rows(path):
    return np.array([[float(x) for x in ln.split()] for ln in open(path) if not ln.lstrip().startswith("#")])


@pytest.mark.parametrize("name", ["SimpleCubic_FM", "SimpleCubic_AFM"])
def test_reference_example_runs_unchanged(tmp_path, name):
    d = tmp_path / name
    shutil.copytree(os.path.join(EX, name), d)
    say(f"example {name}")
    r = subprocess.run([sys.executable, "-m", "ether_amd", str(d)], cwd=ROOT, capture_output=True,
                       text=True, timeout=600)
    say(r.stdout[-300:])
    assert r.returncode == 0, r.stdout + r.stderr
    m = read_directory(str(d))
    mag = _rows(d / "magnetization.dat")
    en = _rows(d / "energy.dat")
    assert mag.shape == (40, 7) and en.shape == (40, 8)
    assert mag[0, 0] == pytest.approx(80.0) and mag[-1, 0] == pytest.approx(2.0)
    T = 2.0
    kT = KB_EV * T
    # |M| and e at the lowest temperature (each repeat equilibrates from a
    # random state, with proposals accepted about 1% of the time at 2 K)
    assert mag[-1, 1] > 0.9
    assert en[-1, 1] == pytest.approx(-0.009 + kT, rel=0.03)
    # ordered at 2 K, disordered at 80 K (kT = 2.3|J|)
    assert mag[0, 1] < 0.3 < mag[-1, 1]
    assert np.all((en[:, 7] > 0) & (en[:, 7] <= 100))
    # moment vectors per species: both near 1 per site and parallel
    mv = _rows(d / "moment_vectors.dat")
    assert mv.shape == (40, 1 + 3 * 2 + 2)
    a, b = mv[-1, 1:4], mv[-1, 4:7]
    assert mv[-1, 7] > 0.9 and mv[-1, 8] > 0.9
    assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.98
    # the direct-access records agree with the text files
    nrec = outputs.TOTAL_OBSERVABLES + 3 * m.n_species
    obs = outputs.read_direct(str(d / "ETHER.obs"), nrec)
    assert np.allclose(obs[:, 1], mag[:, 1], rtol=1e-4, atol=1e-9)
    assert np.allclose(obs[:, 7], en[:, 1], rtol=1e-4, atol=1e-9)
    spn = outputs.read_direct(str(d / "ETHER.spn"), 9 * m.n_sites).reshape(40, m.n_sites, 9)
    assert np.allclose(np.linalg.norm(spn[:, :, 1:4], axis=2), 1.0, atol=1e-6)
    assert os.path.getsize(d / "spiral_state_parameters.dat") == 0
