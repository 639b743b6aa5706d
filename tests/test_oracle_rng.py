"""Oracle pinning: the counter-based generator and the spin proposals.

Philox4x32-10 is pinned by the known-answer vectors of Salmon et al.'s
Random123 distribution.  The proposals (src/George_Marsaglia.f90:29-39,
src/update_spin_details.f90:28-51) are pinned by what the mathematics fixes:
Marsaglia's points are uniform on S^2 (<S>=0, <S_z^2>=1/3, flat S_z histogram),
ANGLE gives S_z ~ U(-1,1) and phi' - phi ~ U(-dphi, dphi), Ising flips S_z.
"""
import math

import numpy as np
import pytest

from oracle import oracle as o

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert list(o.philox(ctr, key)) == out


def test_rng_counter_layout():
    # ctr = {id, replica, step, purpose<<16 | attempt}, key = {seed lo, seed hi}
    seed = 0x0123456789ABCDEF
    a = o.rng(seed, 5, 7, 11, 3, 2)
    b = o.philox([5, 7, 11, (3 << 16) | 2], [0x89ABCDEF, 0x01234567])
    assert list(a) == list(b)


def _one_site_oracle(model=1, angle=False, dphi=0.0):
    return o.Oracle(model, 1, 1, [1], [1.0], [0], [0], [0.0, 0.0, 0.0], [1.0], angle=angle,
                    dphi=dphi)


def test_marsaglia_uniform_on_sphere():
    orc = _one_site_oracle()
    n = 20000
    S = np.array([orc.propose(1992, i, 0, 0, [0, 0, 1.0])[0] for i in range(n)])
    assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-14)
    se = 1.0 / math.sqrt(3 * n)
    assert np.all(np.abs(S.mean(axis=0)) < 5 * se)
    assert abs((S[:, 2] ** 2).mean() - 1.0 / 3.0) < 5 * math.sqrt(4.0 / 45.0 / n)
    # S_z uniform on [-1,1]: 10-bin chi-square
    h, _ = np.histogram(S[:, 2], bins=10, range=(-1, 1))
    chi2 = ((h - n / 10) ** 2 / (n / 10)).sum()
    assert chi2 < 30.0  # 9 dof, p ~ 4e-4


def test_marsaglia_formula_first_candidate():
    # when the first candidate pair is inside the disc, S = (2v1 r, 2v2 r, 1-2s)
    orc = _one_site_oracle()
    for i in range(200):
        w = o.rng(7, i, 0, 0, o.RNG_METROPOLIS, 0)
        a1 = int(w[0] >> 8) - (1 << 23)
        a2 = int(w[1] >> 8) - (1 << 23)
        if a1 * a1 + a2 * a2 >= (1 << 46):
            continue
        v1, v2 = a1 / 2.0 ** 23, a2 / 2.0 ** 23
        s = v1 * v1 + v2 * v2
        exp = np.array([2 * v1 * math.sqrt(1 - s), 2 * v2 * math.sqrt(1 - s), 1 - 2 * s])
        exp /= np.linalg.norm(exp)
        got, phi, eta = orc.propose(7, i, 0, 0, [0, 0, 1.0])
        assert np.allclose(got, exp, atol=1e-15)
        assert phi == 0.0
        assert eta == ((int(w[2]) >> 8) + 1) / 2.0 ** 24


def test_angle_move_distribution():
    dphi = math.radians(5.0)
    orc = _one_site_oracle(angle=True, dphi=dphi)
    n = 20000
    phi0 = 0.3
    sz, dp = [], []
    for i in range(n):
        s, phi, _ = orc.propose(3, i, 0, 0, [1.0, 0, 0], phi_old=phi0)
        sz.append(s[2])
        dp.append(phi - phi0)
        assert abs(np.linalg.norm(s) - 1.0) < 1e-14
        assert abs(math.atan2(s[1], s[0]) - phi) < 1e-9
    sz, dp = np.array(sz), np.array(dp)
    assert np.all(np.abs(dp) <= dphi + 1e-15)
    assert abs(dp.mean()) < 5 * dphi / math.sqrt(3 * n)
    assert abs(sz.mean()) < 5 / math.sqrt(3 * n)
    assert abs((dp ** 2).mean() - dphi ** 2 / 3) < 5 * dphi ** 2 * math.sqrt(4 / 45 / n)


def test_ising_flip_and_threshold():
    orc = _one_site_oracle(model=0)
    s, _, _ = orc.propose(1, 0, 0, 0, [0, 0, 1.0])
    assert list(s) == [0, 0, -1.0]
    # threshold: accept iff w < floor(p 2^32), p >= 1 always
    assert o.ising_threshold(1.0, -1.0) == 0xFFFFFFFF
    assert o.ising_threshold(1.0, 0.0) == 0xFFFFFFFF
    p = math.exp(-2.0)
    assert o.ising_threshold(1.0, 2.0) == math.floor(p * 2 ** 32)
