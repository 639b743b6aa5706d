"""Spiral-state parameter on the device against the oracle.

The measure pass sums S_i x S_j over every site's spiral partners in FP64
from the FP32 spin planes; psi = |sum| / spiral_count must equal the
oracle's FP64 sum over the same spins (read back from the device) to 1e-10,
since only the summation order differs.  Planar spirals pin the absolute
value (psi = |sin k a|).  Sampling passes add psi, psi^2, psi^4 to the
replica's temperature slot (src/get_spiral_state.f90:114-120).
"""
import numpy as np
import pytest

from ether_amd import EtherError
from ether_amd import engine as E
from ether_amd import workloads
from ether_amd.driver import attach_neighbours, make_engine, spiral_partners
from oracle import oracle as o
from tests.helpers import random_xyz, say

pytestmark = pytest.mark.gpu

CASES = [
    ("sc_heis_afm_32", dict(sc="6 6 6", ssp=".T.", ss_dis="1.01", ss_lat="2", ss_direc="0 0 1")),
    ("tri_afm_96", dict(sc="12 12 1", ssp=".T.", ss_dis="1.01", ss_lat="2", ss_direc="1 0 0")),
    ("pyro_pt_16", dict(sc="2 2 2", ssp=".T.", ss_dis="3", ss_lat="2",
                        ss_direc="0 0.7071067811865476 0.7071067811865476")),
]


def _engine(name, over, R, seed=5):
    m = workloads.model(name, **over)
    attach_neighbours(m)
    eng = make_engine(m, R, seed=seed)
    eng.set_betas(m.betas(np.linspace(5, 1, R)))
    pr, cnt = spiral_partners(m)
    assert cnt > 0 and eng.spiral_count == cnt
    return m, eng, pr, cnt


def _set_all(eng, states):
    for r, S in enumerate(states):
        eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))


@pytest.mark.parametrize("name,over", CASES)
def test_psi_matches_oracle_on_random_states(name, over):
    R = 70
    say(f"spiral psi {name} R={R}")
    m, eng, pr, cnt = _engine(name, over, R)
    rng = np.random.default_rng(3)
    _set_all(eng, [random_xyz(rng, m.n_sites) for _ in range(R)])
    eng.measure(False)
    _, psi = eng.spiral()
    for r in range(R):
        S = eng.get_spins(r)[0]
        assert psi[r] == pytest.approx(o.spiral_psi(S, pr, cnt), abs=1e-10), r
    eng.close()


def test_psi_of_planar_spirals():
    name, over = CASES[0]
    R = 8
    say("spiral psi planar")
    m, eng, pr, cnt = _engine(name, over, R)
    ks = np.linspace(0.1, 3.0, R)
    q, _ = np.linalg.qr(np.random.default_rng(9).normal(size=(3, 3)))
    states = []
    for k in ks:
        ph = k * m.positions[:, 2]
        states.append(np.stack([np.cos(ph), np.sin(ph), 0 * ph], axis=1) @ q.T)
    _set_all(eng, states)
    eng.measure(False)
    _, psi = eng.spiral()
    assert np.allclose(psi, np.abs(np.sin(ks)), atol=2e-6)
    eng.close()


def test_sampled_moments_and_reset():
    name, over = CASES[1]
    R = 12
    say("spiral moments")
    m, eng, pr, cnt = _engine(name, over, R)
    rng = np.random.default_rng(4)
    want = np.zeros((R, 3))
    for _ in range(3):
        states = [random_xyz(rng, m.n_sites) for _ in range(R)]
        _set_all(eng, states)
        eng.measure(True)
        for r in range(R):
            p = o.spiral_psi(eng.get_spins(r)[0], pr, cnt)
            want[r] += [p, p * p, p**4]
    sums, _ = eng.spiral()
    assert np.allclose(sums, want, rtol=1e-9, atol=1e-14)
    assert np.all(eng.moments()[:, E.MOM_N] == 3)
    eng.reset_moments()
    sums, _ = eng.spiral()
    assert np.all(sums == 0)
    eng.close()


def test_spiral_through_tempering_run():
    name, over = CASES[2]
    R = 16
    say("spiral under PT")
    m, eng, pr, cnt = _engine(name, over, R)
    eng.set_ladders([0, 8, 16])
    rng = np.random.default_rng(6)
    _set_all(eng, [random_xyz(rng, m.n_sites) for _ in range(R)])
    # swaps every 3 MCS except on sampling steps (src/ether.f90:189-198)
    sc = eng.run(1, 60, 10, 5, 0, 0.0, 3, 0)
    assert sc == 16
    mom = eng.moments()
    sums, psi = eng.spiral()
    n = mom[:, E.MOM_N]
    assert np.all(n == 10)
    p, p2, p4 = sums[:, 0] / n, sums[:, 1] / n, sums[:, 2] / n
    assert np.all((p > 0) & (p <= 1)) and np.all(p2 <= p + 1e-15) and np.all(p4 <= p2 + 1e-15)
    assert np.all(p2 >= p * p * (1 - 1e-12))
    # psi of the last measured state belongs to the replica, not the slot
    for r in range(R):
        assert psi[r] == pytest.approx(o.spiral_psi(eng.get_spins(r)[0], pr, cnt), abs=1e-10)
    eng.close()


def test_ising_refuses_spiral():
    m = workloads.model("sc_ising_16", sc="4 4 4")
    attach_neighbours(m)
    eng = E.Engine(m.model_id, m.n_sites, 4, 1)
    with pytest.raises(EtherError) as e:
        eng.set_spiral(np.full((m.n_sites, 2), -1, np.int32))
    assert e.value.code == E.ETHER_ERR_ARG
    eng.close()
