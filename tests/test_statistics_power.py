"""The whole-run agreement criterion (helpers.assert_runs_agree) on the CPU:
two oracle runs of one model with different seeds pass it, and the same with
a mistake planted in one oracle run (every temperature 3% high, as a wrong
k_B or unit would give) fails it.  The oracle is the reference's chain
restated (src/Metropolis.f90, src/ether.f90); 6^3 Heisenberg antiferromagnet,
OVRR every MCS, 6 temperatures, 8 repeats."""
import math

import numpy as np
import pytest

from ether_amd import driver, workloads
from ether_amd.inputs import fortran_nint
from oracle import oracle as o
from tests.helpers import assert_runs_agree


def _oracle_run(m, seed, t_scale=1.0):
    t = m.tags
    temps = np.asarray(t.temperatures()) * t_scale
    betas = m.betas(temps)
    driver.attach_neighbours(m)
    orc = o.Oracle.from_model(m)
    n = m.n_sites
    order = np.arange(n, dtype=np.int32)
    per = np.zeros((t.repeat, len(temps), 6))
    for k in range(t.repeat):
        for j, b in enumerate(betas):
            rng = np.random.default_rng([seed, k, j])
            v = rng.normal(size=(n, 3))
            S = np.ascontiguousarray(v / np.linalg.norm(v, axis=1)[:, None])
            mom = orc.run_chain(S, None, order, order, b, seed, k * len(temps) + j, t.tmcs, t.tmcs_eq,
                                fortran_nint(t.coa), t.ovrr_mcs, t.ovrr_para)
            per[k, j] = driver.repeat_observables(mom, b, n)
    return per.mean(axis=0), per.std(axis=0, ddof=1) / math.sqrt(t.repeat)


MODEL = dict(sc="6 6 6", mcs="1000 400", repeat="8", temp="60 10 10")


def test_two_seeds_agree():
    m = workloads.model("sc_heis_afm_32", **MODEL)
    a = _oracle_run(m, 11)
    b = _oracle_run(m, 12)
    assert_runs_agree(a[0], a[1], b[0], b[1], 8, 8, "oracle vs oracle")


def test_a_planted_temperature_error_is_caught():
    m = workloads.model("sc_heis_afm_32", **MODEL)
    a = _oracle_run(m, 11)
    b = _oracle_run(m, 12, t_scale=1.03)
    with pytest.raises(AssertionError):
        assert_runs_agree(a[0], a[1], b[0], b[1], 8, 8, "oracle vs oracle with T 3% high")
