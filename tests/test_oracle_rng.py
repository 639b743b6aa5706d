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


def test_pooled_marsaglia_rule():
    """The device pools Marsaglia candidates over 64 replicas of a site: every
    replica still gets a point uniform on S^2, replicas whose first pair is in
    the disc keep it, and a group of one replica is the plain rule."""
    orc = _one_site_oracle()
    # group of one == plain per-replica proposal
    for i in range(50):
        sn, _, eta = orc.propose_site(11, i, 0, [[0, 0, 1.0]])
        s1, _, eta1 = orc.propose(11, i, 0, 0, [0, 0, 1.0])
        assert np.allclose(sn[0], s1, atol=0) and eta[0] == eta1
    # first pair inside -> identical to the plain rule for that replica
    R = 200
    pts = []
    for i in range(300):
        sn, _, eta = orc.propose_site(5, i, 0, np.tile([0, 0, 1.0], (R, 1)))
        for r in range(R):
            w = o.rng(5, i, r, 0, o.RNG_METROPOLIS, 0)
            a1, a2 = int(w[0] >> 8) - (1 << 23), int(w[1] >> 8) - (1 << 23)
            if a1 * a1 + a2 * a2 < (1 << 46):
                s1, _, _ = orc.propose(5, i, r, 0, [0, 0, 1.0])
                assert np.allclose(sn[r], s1, atol=0)
        pts.append(sn)
    S = np.concatenate(pts)
    n = len(S)
    assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-14)
    assert np.all(np.abs(S.mean(axis=0)) < 5 / math.sqrt(3 * n))
    assert abs((S[:, 2] ** 2).mean() - 1 / 3) < 5 * math.sqrt(4 / 45 / n)
    h, _ = np.histogram(S[:, 2], bins=10, range=(-1, 1))
    assert ((h - n / 10) ** 2 / (n / 10)).sum() < 30.0


def test_batch_sweep_matches_single_replica_sweeps_for_ising():
    """Ising draws do not pool: the batch sweep equals per-replica sweeps."""
    from ether_amd import build_neighbours
    cnt, idx = build_neighbours(np.eye(3), [[0, 0, 0]], (4, 4, 4), (0, 0, 0), [1.0])
    orc = o.Oracle(0, 64, 1, np.ones(64, np.int32), np.ones(64), cnt, idx, [0, 0, -1.0], [1.0])
    rng = np.random.default_rng(0)
    R = 9
    S = np.zeros((R, 64, 3))
    S[:, :, 2] = np.where(rng.random((R, 64)) < 0.5, 1.0, -1.0)
    S2 = S.copy()
    order = np.arange(64, dtype=np.int32)
    betas = np.linspace(0.2, 0.6, R)
    orc.sweep_batch(S, None, order, order, betas, 3, 7)
    for r in range(R):
        Sr = np.ascontiguousarray(S2[r])
        orc.sweep(Sr, None, order, order, betas[r], 3, r, 7)
        assert np.array_equal(Sr, S[r])


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
