"""The reference's result files written by ether_amd.outputs.

Closed forms pin the content: in a ferromagnetic state every site's moment
is s along the spin axis, so moment_vectors.dat holds (0, 0, s) and |M| = s
per species; with the checkerboard stagger a Neel state gives the same.
ETHER.obs / ETHER.spn are direct-access files of equal FP64 records
(src/rw_file.f90:35-55): the file is nscan * recl bytes and record t is
the t-th block.  gss.dat is read back the way the reference's utilities
read it (list-directed, utility/Structure_Factor/structure_factor.f90:78-140).
The byte layout of every text record is checked against the Fortran
driver's own formatted output in tests/test_gpu_driver.py.
"""
import os

import numpy as np

from ether_amd import driver, outputs, workloads


def _results(m, S, temps, spiral=False):
    nscan = len(temps)
    rng = np.random.default_rng(1)
    res = driver.Results(temps=np.asarray(temps), mean=rng.random((nscan, 6)), sem=rng.random((nscan, 6)),
                         accept_pct=rng.random(nscan) * 100, per_repeat=None, wall_s=1.0, updates=1.0,
                         final_S=S, final_phi=np.zeros(S.shape[:2]), model=m)
    if spiral:
        res.spiral_mean, res.spiral_sem = rng.random((nscan, 3)), rng.random((nscan, 3))
    return res


def test_files_and_records(tmp_path):
    m = workloads.model("sc_heis_afm_32", sc="4 4 4", temp="60 6 27", cb=".T.")
    temps = m.tags.temperatures()
    nscan, N = len(temps), m.n_sites
    # t = 0: Neel state along z (stagger * S = +z everywhere); others random
    rng = np.random.default_rng(2)
    S = rng.normal(size=(nscan, N, 3))
    S /= np.linalg.norm(S, axis=2, keepdims=True)
    S[0] = 0.0
    S[0, :, 2] = m.stagger
    res = _results(m, S, temps, spiral=True)
    res.write(str(tmp_path))
    names = set(os.listdir(tmp_path))
    for f in ["ETHER.obs", "ETHER.spn", "gss.dat", "moment_vectors.dat", "energy.dat",
              "magnetization.dat", "spiral_state_parameters.dat"] + [outputs.spk_name(T) for T in temps]:
        assert f in names, f
    # ETHER.obs: nscan records of 20 + 3 nspecies doubles in process_observables order
    nrec = outputs.TOTAL_OBSERVABLES + 3 * m.n_species
    assert os.path.getsize(tmp_path / "ETHER.obs") == nscan * nrec * 8
    obs = outputs.read_direct(str(tmp_path / "ETHER.obs"), nrec)
    assert np.array_equal(obs[:, 0], temps)
    assert np.array_equal(obs[:, 1], res.mean[:, 3]) and np.array_equal(obs[:, 8], res.mean[:, 2])
    assert np.array_equal(obs[:, 13], res.accept_pct)
    assert np.array_equal(obs[:, 14], res.spiral_mean[:, 0]) and np.array_equal(obs[:, 18], res.spiral_mean[:, 1])
    assert np.allclose(obs[0, 20:23], [0, 0, N])             # Neel state: M_stag = N s z
    # ETHER.spn: ion(0:8, 1:N) per record
    assert os.path.getsize(tmp_path / "ETHER.spn") == nscan * N * 9 * 8
    spn = outputs.read_direct(str(tmp_path / "ETHER.spn"), 9 * N).reshape(nscan, N, 9)
    assert np.array_equal(spn[:, :, 0], np.tile(np.arange(1, N + 1), (nscan, 1)))
    assert np.array_equal(spn[:, :, 1:4], S)
    assert np.array_equal(spn[2, :, 6:9], m.positions) and np.all(spn[:, :, 4] == 1)
    # moment_vectors.dat row of the Neel state: (0, 0, 1) and |M| = 1 per site
    rows = [ln for ln in open(tmp_path / "moment_vectors.dat") if not ln.lstrip().startswith("#")]
    assert [float(x) for x in rows[0].split()] == [temps[0], 0.0, 0.0, 1.0, 1.0]
    assert len(rows) == nscan


def test_gss_reads_back_like_the_utilities(tmp_path):
    m = workloads.model("pyro_pt_16", sc="2 2 2", temp="1 0.05 0.19")
    temps = m.tags.temperatures()
    rng = np.random.default_rng(3)
    S = rng.normal(size=(len(temps), m.n_sites, 3))
    S /= np.linalg.norm(S, axis=2, keepdims=True)
    _results(m, S, temps).write(str(tmp_path))
    tok = open(tmp_path / "gss.dat").read().split("\n")
    nscan, *sc = [int(x) for x in tok[0].split()]
    assert nscan == len(temps) and sc == [2, 2, 2]
    assert [int(x) for x in tok[1].split()] == [1, 1, 1, 2, 2, 2]
    nb, nsp, ht, lt, tint = tok[2].split()
    assert (int(nb), int(nsp)) == (16, 1) and float(ht) == 1.0 and float(tint) == 0.19
    assert tok[3].strip() == "K" and len(tok[3]) == 20
    assert int(tok[4]) == m.n_sites
    abc = np.array([[float(x) for x in tok[5 + i].split()] for i in range(3)])
    assert np.allclose(abc, m.structure.abc, atol=5e-6)
    pts = np.array([[float(x) for x in ln.split()] for ln in tok[9:9 + m.n_sites]])
    assert np.allclose(pts[:, :3], m.positions, atol=5e-7) and np.array_equal(pts[:, 3], np.arange(1, m.n_sites + 1))
    at = 9 + m.n_sites
    for k, T in enumerate(temps):
        assert float(tok[at]) == round(T, 5)
        blk = np.array([[float(x) for x in ln.split()] for ln in tok[at + 1:at + 1 + m.n_sites]])
        assert np.allclose(blk[:, :3], S[k], atol=5e-7) and np.all(blk[:, 3] == 1)
        at += 1 + m.n_sites


def test_final_state_per_temperature_follows_the_slots():
    class Fake:
        N = 2

        def pt_state(self):
            # 2 temperatures x 2 repeats; the last repeat's slots 2, 3 are held by replicas 3, 2
            return np.array([0, 1, 3, 2]), None, None, None

        def get_spins(self, r):
            return np.full((2, 3), float(r)), np.zeros(2)

    S, _ = outputs.final_states(Fake(), 2, 2)
    assert S[0, 0, 0] == 3.0 and S[1, 0, 0] == 2.0


def test_spk_name_and_body():
    m = workloads.model("sc_heis_afm_32", sc="2 2 2")
    assert outputs.spk_name(60.0) == "60.0000spK.xsf"
    assert outputs.spk_name(0.629032) == "0.6290spK.xsf"
    S = np.tile([0.0, 0.0, 1.0], (m.n_sites, 1))
    lines = outputs.spk_text(m, S).split("\n")
    assert lines[0] == " CRYSTAL" and lines[9] == " PRIMCOORD"
    assert [float(x) for x in lines[2].split()] == [2.0, 0.0, 0.0]
    body = lines[11]
    assert body[:5] == "   A " and len(body) == 5 + 6 * 14      # A5 of the character(len=2) name
    assert [float(x) for x in body[5:].split()] == [0.0, 0.0, 0.0, 0.0, 0.0, 1.0]
