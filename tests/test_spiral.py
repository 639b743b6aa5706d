"""Spiral-state parameter (ssp): partner search and estimators.

Partners (src/get_spiral_state.f90:47-82): the oracle restates the
reference's O(N^2) scan; it is pinned by closed forms on cubic grids (with
dir = z and a small cone every site below the top layer has exactly its +z
neighbour; with a diagonal dir, exactly its (+1,+1,0) neighbour).  The
library's grid-binned builder (ether_build_spiral_partners, host code) must
return the oracle's lists entry for entry, on grids, random clouds and the
pyrochlore supercell.

psi (:84-120) is pinned by rotated planar spirals: with S_i =
Q (cos k z_i, sin k z_i, 0) every +z pair contributes Q (0, 0, sin k a), so
psi = sin(k a) whatever the rotation Q, and oppositely wound halves cancel.
The estimators (:122-140) are pinned on constant and two-valued series.
"""
import numpy as np
import pytest

from ether_amd import EtherError, build_spiral_partners, workloads
from oracle import oracle as o


def _grid(L, a=1.0):
    z, y, x = np.meshgrid(np.arange(L[2]), np.arange(L[1]), np.arange(L[0]), indexing="ij")
    return a * np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64)


def _rot(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q


def test_oracle_partners_closed_forms():
    L = (4, 5, 6)
    pos = _grid(L)
    n = pos.shape[0]
    pr, cnt = o.spiral_partners(pos, (0, 0, 1), 1.01, 2.0, 5)
    assert cnt == n - L[0] * L[1]
    for i in range(n):
        x, y, z = pos[i].astype(int)
        want = [i + L[0] * L[1]] if z < L[2] - 1 else []
        got = [j for j in pr[i] if j >= 0]
        assert got == want, i
    # diagonal direction: only (+1, +1, 0) within 1.5 and 2 degrees
    pr, cnt = o.spiral_partners(pos, np.array([1, 1, 0]) / np.sqrt(2), 1.5, 2.0, 5)
    assert cnt == (L[0] - 1) * (L[1] - 1) * L[2]
    i = 0
    assert [j for j in pr[i] if j >= 0] == [1 + L[0]]
    # a wide cone around z picks the 5 sites of the upper hemisphere within sqrt(2)
    # that make at most 45 degrees with z: (0,0,1) and the 4 (+-1,0,1),(0,+-1,1)
    pr, cnt = o.spiral_partners(pos, (0, 0, 1), 1.5, 45.0 + 1e-9, 5)
    inner = sum(1 for p in pos if 0 < p[0] < L[0] - 1 and 0 < p[1] < L[1] - 1 and p[2] < L[2] - 1)
    full = sum(1 for i in range(n) if (pr[i] >= 0).sum() == 5)
    assert full == inner
    # overflow: more partners than the capacity
    _, cnt = o.spiral_partners(pos, (0, 0, 1), 1.5, 90.0, 2)
    assert cnt == -1


@pytest.mark.parametrize("case", ["grid_z", "grid_diag", "cloud", "cloud_tilted", "pyrochlore"])
def test_builder_matches_oracle(case):
    rng = np.random.default_rng(7)
    if case == "grid_z":
        pos, d, dis, lat, cap = _grid((6, 6, 6)), (0, 0, 1), 1.01, 2.0, 5
    elif case == "grid_diag":
        pos, d, dis, lat, cap = _grid((5, 6, 7), 2.5), np.array([1, 1, 1]) / np.sqrt(3), 4.4, 2.0, 5
    elif case == "cloud":
        pos, d, dis, lat, cap = rng.uniform(0, 20, (700, 3)), (0, 0, 1), 2.5, 40.0, 16
    elif case == "cloud_tilted":
        v = rng.normal(size=3)
        pos, d, dis, lat, cap = rng.uniform(-5, 30, (900, 3)), v / np.linalg.norm(v), 3.0, 60.0, 24
    else:
        m = workloads.model("pyro_pt_16", sc="3 3 3")
        pos, d, dis, lat, cap = m.positions, (0, 0, 1), 4.0, 2.0, 5   # the reference's defaults
    p1, c1 = build_spiral_partners(pos, d, dis, lat, cap)
    p2, c2 = o.spiral_partners(pos, d, dis, lat, cap)
    assert c2 >= 0 and c1 == c2
    assert np.array_equal(p1, p2)
    if case != "pyrochlore":
        assert c1 > 0


def test_builder_overflow_is_an_error():
    with pytest.raises(EtherError) as e:
        build_spiral_partners(_grid((4, 4, 4)), (0, 0, 1), 1.5, 90.0, 2)
    assert e.value.code == 4   # ETHER_ERR_LIMIT


def _spiral_state(pos, k, Q, wind=None):
    ph = k * pos[:, 2] * (1.0 if wind is None else wind)
    S = np.stack([np.cos(ph), np.sin(ph), np.zeros_like(ph)], axis=1)
    return S @ Q.T


@pytest.mark.parametrize("k", [0.0, 0.3, np.pi / 2, 2.0])
def test_psi_of_planar_spiral(k):
    L = (4, 4, 8)
    pos = _grid(L)
    pr, cnt = o.spiral_partners(pos, (0, 0, 1), 1.01, 2.0, 5)
    S = _spiral_state(pos, k, _rot(3))
    assert o.spiral_psi(S, pr, cnt) == pytest.approx(abs(np.sin(k)), abs=1e-13)
    # left half wound one way, right half the other: the chiralities cancel
    wind = np.where(pos[:, 0] < L[0] / 2, 1.0, -1.0)
    S2 = _spiral_state(pos, k, _rot(4), wind)
    assert o.spiral_psi(S2, pr, cnt) == pytest.approx(0.0, abs=1e-13)


def test_psi_is_brute_force_cross_sum():
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 6, (60, 3))
    pr, cnt = o.spiral_partners(pos, (0, 1, 0), 3.0, 70.0, 60)
    S = rng.normal(size=(60, 3))
    S /= np.linalg.norm(S, axis=1, keepdims=True)
    tot = np.zeros(3)
    for i in range(60):
        for j in pr[i]:
            if j >= 0:
                tot += np.cross(S[i], S[j])
    assert cnt > 0
    assert o.spiral_psi(S, pr, cnt) == pytest.approx(np.linalg.norm(tot) / cnt, rel=1e-12)


def test_spiral_estimators():
    # constant psi = c: U = 1.5 - 0.5 = 1, chi = 0
    c, n = 0.37, 50
    out = o.spiral_observables([n * c, n * c**2, n * c**4], n, 2.0, 100)
    assert out == pytest.approx([c, 1.0, 0.0], abs=1e-14)
    # two values a, b half the time each
    a, b, beta, cnt = 0.2, 0.6, 3.0, 77
    sums = [n * (a + b), n * (a * a + b * b), n * (a**4 + b**4)]
    out = o.spiral_observables(sums, 2 * n, beta, cnt)
    p2 = (a * a + b * b) / 2
    assert out[0] == pytest.approx((a + b) / 2)
    assert out[1] == pytest.approx(1.5 - 0.5 * ((a**4 + b**4) / 2) / p2**2)
    assert out[2] == pytest.approx(beta * ((a - b) / 2) ** 2 * cnt)
