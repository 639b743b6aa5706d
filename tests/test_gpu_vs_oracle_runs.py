"""Distributional parity of whole runs: GPU against the CPU oracle.

The GPU run goes through the batched driver: colour-ordered sweeps, every
temperature x repeat chain on the device, and the schedule of src/ether.f90.
The oracle runs each chain on its own, in the reference's typewriter order
(src/Metropolis.f90:39) and with an independent seed.  Both start from the
driver's start state.  Per temperature, the reference's estimators
(e, U_e, C_v, m, U_m, chi with their SEM over repeats,
src/observables_per_repeat.f90) must agree by Welch's t on their means
(helpers.assert_runs_agree): the smallest p-value above 0.001 / K for K
comparisons, and no more p < 0.01 than chance allows.  The benchmark
lattices at full size are in test_gpu_statistics_sizes.py.

The reference itself cannot be built in this image (its `use mpi` needs an
MPI module in a format the image's Fortran compiler reads, and there is
none), so this is the 2-sigma parity check of the sampled distribution, run
against the oracle's restatement of the reference's chain.  With OVRR on, the
GPU runs the reflection fused into each colour phase while the oracle runs
the reference's whole-pass order: the check covers that reordering too.
"""
import numpy as np
import pytest

from ether_amd import driver, workloads
from ether_amd.inputs import fortran_nint
from oracle import oracle as o
from tests.helpers import assert_runs_agree, say

pytestmark = pytest.mark.gpu


def _oracle_run(m, seed):
    t = m.tags
    temps = t.temperatures()
    betas = m.betas(temps)
    orc = o.Oracle.from_model(m)
    n = m.n_sites
    order = np.arange(n, dtype=np.int32)
    coa = fortran_nint(t.coa)
    ovrr_on = t.ovrr and t.xyz and not t.zeeman and not t.sia
    rng = np.random.default_rng(seed)
    per = np.zeros((t.repeat, len(temps), 6))
    for k in range(t.repeat):
        for j, b in enumerate(betas):
            # the driver's start: fresh_spins applied to the generate_supercell state
            if t.ising:
                S = np.ascontiguousarray(-m.initial_spins)       # Ising: every spin flipped
                phi = None
            elif t.angle:
                # ANGLE: fresh S_z, phi = phi_basis + dphi * U(-1,1)
                sz = rng.uniform(-1, 1, n)
                phi = np.ascontiguousarray(m.initial_phi + t.dphi * rng.uniform(-1, 1, n))
                rt = np.sqrt(1 - sz * sz)
                S = np.ascontiguousarray(np.stack([rt * np.cos(phi), rt * np.sin(phi), sz], axis=1))
            else:
                v = rng.normal(size=(n, 3))                     # Marsaglia: uniform on S^2
                S = np.ascontiguousarray(v / np.linalg.norm(v, axis=1)[:, None])
                phi = None
            mom = orc.run_chain(S, phi, order, order, b, seed, k * len(temps) + j, t.tmcs,
                                t.tmcs_eq, coa, t.ovrr_mcs if ovrr_on else 0,
                                t.ovrr_para if ovrr_on else 0.0)
            per[k, j] = driver.repeat_observables(mom, b, n)
    mean = per.mean(axis=0)
    sem = per.std(axis=0, ddof=1) / np.sqrt(t.repeat)
    return mean, sem


@pytest.mark.parametrize("name,over", [
    ("sc_ising_16", dict(sc="6 6 6", mcs="1500 500", repeat="8", temp="100 10 10")),
    ("sc_heis_afm_32", dict(sc="6 6 6", mcs="1000 400", repeat="8", temp="60 6 6")),
    # SIA + Zeeman + ANGLE moves (30 degree window so that 1500 MCS equilibrate)
    ("tri_afm_96", dict(sc="6 6 1", mcs="4000 1500", repeat="8", temp="20 2 2", angle=".T. 30")),
])
def test_observables_agree_within_two_sigma(name, over):
    say(f"runs {name}")
    m = workloads.model(name, **over)
    res = driver.run(m)
    mean_o, sem_o = _oracle_run(m, 777)
    assert_runs_agree(res.mean, res.sem, mean_o, sem_o, m.tags.repeat, m.tags.repeat, name)
