"""fresh_spins (src/fresh_spins.f90:31-38) on the GPU: ETHER_INIT_FRESH gives
every site of every replica one unconditional proposal
(src/update_spin_details.f90:20-53).

  - Ising: every spin is flipped.
  - XYZ Marsaglia: each site gets the oracle's o_propose point for the init
    draw (purpose RNG_INIT, step = the init call count) within FP32 rounding,
    and over R x N sites the points are uniform on S^2: <S> = 0,
    <S_z^2> = 1/3, flat S_z histogram.
  - ANGLE: S_z ~ U(-1,1) flat, phi' - phi_old within [-delta, delta], and
    each site matches the oracle within FP32 rounding.
"""
import math

import numpy as np
import pytest

from ether_amd import engine as E
from tests.helpers import colour_order, pair, random_xyz, say
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def test_init_fresh_ising_flips_every_spin():
    say("init fresh ising")
    R = 70
    m, eng, _ = pair("sc_ising_16", R, sc="6 6 6")
    eng.set_betas(m.betas(np.linspace(10, 100, R)))
    rng = np.random.default_rng(3)
    states = []
    for r in range(R):
        S = np.zeros((m.n_sites, 3))
        S[:, 2] = np.where(rng.random(m.n_sites) < 0.5, 1.0, -1.0)
        eng.set_spins(r, S)
        states.append(S)
    eng.init_spins(E.INIT_FRESH)
    for r in range(R):
        Sg, _ = eng.get_spins(r)
        assert np.array_equal(Sg[:, 2], -states[r][:, 2]), f"replica {r}"
    eng.close()


@pytest.mark.parametrize("angle", [False, True])
def test_init_fresh_xyz_matches_oracle_and_is_uniform(angle):
    say(f"init fresh xyz angle={angle}")
    R = 64
    over = dict(sc="12 12 1", sia=".F.", zeeman=".F. 0 0 0")
    if not angle:
        over["angle"] = ".F. 0"
    m, eng, orc = pair("tri_afm_96", R, **over)
    assert bool(m.tags.angle) == angle
    eng.set_betas(m.betas(np.linspace(1, 20, R)))
    rng = np.random.default_rng(5)
    old = []
    for r in range(R):
        S = random_xyz(rng, m.n_sites).astype(np.float32).astype(np.float64)
        phi = np.arctan2(S[:, 1], S[:, 0]).astype(np.float32).astype(np.float64)
        eng.set_spins(r, S, phi)
        old.append((S, phi))
    # two calls: the second draws with step 1 from the state the first left
    _, _, _, rid = colour_order(eng)
    for call in range(2):
        eng.init_spins(E.INIT_FRESH)
        new = [eng.get_spins(r) for r in range(R)]
        worst = 0.0
        for r in (0, 17, R - 1):
            S0, p0 = old[r]
            for i in range(0, m.n_sites, 7):
                sn, pn, _ = orc.propose(m.tags.seed, int(rid[i]), r, call, S0[i], p0[i],
                                        purpose=o.RNG_INIT)
                worst = max(worst, float(np.abs(new[r][0][i] - sn).max()))
                if angle:
                    assert abs(math.remainder(new[r][1][i] - pn, 2 * math.pi)) < 1e-5
        assert worst < 2e-6, worst
        Z = np.concatenate([S[:, 2] for S, _ in new])
        allS = np.concatenate([S for S, _ in new])
        n = Z.size
        assert np.allclose(np.linalg.norm(allS, axis=1), 1.0, atol=2e-6)
        h, _ = np.histogram(Z, bins=20, range=(-1, 1))
        assert ((h - n / 20) ** 2 / (n / 20)).sum() < 45.0
        assert abs((Z ** 2).mean() - 1 / 3) < 4.5 * math.sqrt(4 / 45 / n)
        if angle:
            dphi = m.tags.dphi
            for r in range(R):
                d = np.remainder(new[r][1] - old[r][1] + math.pi, 2 * math.pi) - math.pi
                assert np.all(np.abs(d) <= dphi * (1 + 1e-5) + 1e-6)
        else:
            assert np.all(np.abs(allS.mean(axis=0)) < 4.5 / math.sqrt(3 * n))
        old = [(S.copy(), p.copy()) for S, p in new]
    eng.close()
