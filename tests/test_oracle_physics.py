"""Oracle pinning: energies, sweeps, overrelaxation, estimators, tempering.

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign, a transposed index, the Zeeman/SIA counting of
src/get_tot_energy.f90:40) fails it:
  - closed forms: FM/Neel simple cubic, 120-degree triangular, a pyrochlore
    state with zero moment on every tetrahedron;
  - an O(N^2) pair sum written from positions (not from the neighbour lists)
    against o_total_energy and o_delta_energy;
  - exact enumeration of the 4x4 Ising model against the oracle's Metropolis;
  - a single spin in a field (Langevin function) and with uniaxial SIA;
  - overrelaxation: |S'| = 1, S'.A = S.A, reflection twice = identity.
"""
import itertools
import math

import numpy as np
import pytest

from ether_amd import build_neighbours
from ether_amd import workloads
from ether_amd.driver import over_repeats, repeat_observables
from oracle import oracle as o


def sc_oracle(L, J, model=1, shells=(1.0,), Lz=None, **kw):
    """Simple cubic (or square when Lz=1 with c=10) lattice, 1 species, J per shell."""
    Lz = L if Lz is None else Lz
    abc = np.diag([1.0, 1.0, 1.0 if Lz > 1 else 10.0])
    cnt, idx = build_neighbours(abc, [[0, 0, 0]], (L, L, Lz), (0, 0, 0), list(shells))
    n = L * L * Lz
    jx = np.zeros((len(shells), 1, 1, 3))
    for s, j in enumerate(J if np.ndim(J) else [J]):
        jx[s, 0, 0] = j
    orc = o.Oracle(model, n, 1, np.ones(n, np.int32), np.ones(n), cnt, idx, jx.reshape(-1), [1.0], **kw)
    return orc, cnt, idx


def random_spins(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def test_fm_and_neel_simple_cubic():
    L = 4
    orc, _, _ = sc_oracle(L, -1.0)
    S = np.tile([0, 0, 1.0], (L ** 3, 1))
    ex, ze, si = orc.total_energy(S)
    assert ex / (2 * L ** 3) == pytest.approx(3 * -1.0)      # e = 3J
    assert ze == 0.0 and si == 0.0
    # single flip in the all-up state: dE = -2 sum_j J = -12 J
    assert orc.delta_energy(S, 5, [0, 0, -1.0]) == pytest.approx(12.0)
    orc2, _, _ = sc_oracle(L, 1.0)
    x, y, z = np.meshgrid(range(L), range(L), range(L), indexing="ij")
    sign = np.where(((x + y + z) % 2) == 0, 1.0, -1.0).transpose(2, 1, 0).ravel()
    S = np.zeros((L ** 3, 3))
    S[:, 2] = sign
    assert orc2.total_energy(S)[0] / (2 * L ** 3) == pytest.approx(-3.0)


def test_triangular_120_degree_state():
    m = workloads.model("tri_afm_96", sc="6 6 1", zeeman=".F. 0 0 0", sia=".F.", para=".T. 1")
    cnt, idx = build_neighbours(m.structure.abc, m.basis_xyz, m.sc, (0, 0, 0), m.shells)
    orc = o.Oracle(1, m.n_sites, 1, m.species, m.stagger, cnt, idx, m.j_exc.reshape(-1), m.spin)
    # sublattice (i + 2j) mod 3 for a1=(1,0), a2=(1/2, sqrt3/2) gives the 120-degree state
    n = m.n_sites
    i = np.arange(n) % 6
    j = np.arange(n) // 6
    ang = 2 * np.pi / 3 * ((i + 2 * j) % 3)
    S = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], axis=1)
    assert orc.total_energy(S)[0] / (2 * n) == pytest.approx(-1.5)


def test_pyrochlore_two_in_two_out_energy():
    m = workloads.model("pyro_pt_16", sc="2 2 2", para=".T. 1")
    cnt, idx = build_neighbours(m.structure.abc, m.basis_xyz, m.sc, (0, 0, 0), m.shells[:1])
    jx = m.j_exc[:1].reshape(-1)
    orc = o.Oracle(1, m.n_sites, 1, m.species, m.stagger, cnt, idx, jx, m.spin)
    frac = np.round(np.linalg.solve(m.structure.abc.T, m.basis_xyz.T).T * 4).astype(int) % 2
    sub = [tuple(f) for f in frac]
    up = {(0, 0, 0), (1, 1, 0)}
    basis_sign = np.array([1.0 if s in up else -1.0 for s in sub])
    assert sorted(set(sub)) == sorted({(0, 0, 0), (1, 1, 0), (1, 0, 1), (0, 1, 1)})
    nb = len(basis_sign)
    S = np.zeros((m.n_sites, 3))
    S[:, 2] = basis_sign[np.arange(m.n_sites) % nb]
    ex = orc.total_energy(S)[0]
    assert ex / 2 / m.n_sites == pytest.approx(-1.0)        # E_phys/N = -J


def _pair_sum_energy(pos, abc_sc, species, shells, jtab, S, h=None, g=1.0, mub=1.0, spin=None,
                     sia=None, tol=1e-5):
    """Independent O(N^2) physical energy from positions: 1/2 sum_ij J o S_i S_j + Zeeman + SIA."""
    n = len(pos)
    e = 0.0
    shifts = [np.array(s) for s in itertools.product([0, 1, -1], repeat=3)]
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            for s, d in enumerate(shells):
                hit = any(abs(d - np.linalg.norm(pos[i] - (pos[j] + sh @ abc_sc))) <= tol
                          for sh in shifts)
                if hit:
                    e += 0.5 * np.dot(jtab[s, species[i] - 1, species[j] - 1], S[i] * S[j])
    for i in range(n):
        if h is not None:
            e += -g * mub * spin[species[i] - 1] * np.dot(S[i], h)
        if sia is not None:
            e += np.dot(sia[species[i] - 1], S[i] ** 2)
    return e


def test_energy_against_independent_pair_sum():
    rng = np.random.default_rng(5)
    L = 3
    abc = np.eye(3)
    shells = [1.0, 2 ** 0.5, 3 ** 0.5]
    jt = rng.normal(size=(3, 2, 2, 3))
    jt = 0.5 * (jt + jt.transpose(0, 2, 1, 3))
    basis = [[0, 0, 0], [0.5, 0.5, 0.5]]
    shells = [1.0, 2 ** 0.5, 3 ** 0.5, 0.75 ** 0.5]
    jt = rng.normal(size=(4, 2, 2, 3))
    jt = 0.5 * (jt + jt.transpose(0, 2, 1, 3))
    cnt, idx = build_neighbours(abc, basis, (L, L, L), (0, 0, 0), shells)
    n = 2 * L ** 3
    species = np.tile([1, 2], L ** 3).astype(np.int32)
    pos = np.array([[i, j, k] for k in range(L) for j in range(L) for i in range(L)], float)
    pos = (pos[:, None, :] + np.array(basis)[None]).reshape(-1, 3)
    h = np.array([0.3, -0.2, 0.7])
    spin = np.array([1.0, 2.5])
    sia = rng.normal(size=(2, 3))
    orc = o.Oracle(1, n, 2, species, np.ones(n), cnt, idx, jt.reshape(-1), spin, zeeman=h,
                   g_factor=2.0, mu_b=0.4, sia=sia.reshape(-1))
    S = random_spins(rng, n)
    ref = _pair_sum_energy(pos, abc * L, species, shells, jt, S, h, 2.0, 0.4, spin, sia)
    ex, ze, si = orc.total_energy(S)
    assert 0.5 * ex + ze + si == pytest.approx(ref, rel=1e-12, abs=1e-12)
    for i in (0, 7, 31, n - 1):
        sn = random_spins(rng, 1)[0]
        S2 = S.copy()
        S2[i] = sn
        ref2 = _pair_sum_energy(pos, abc * L, species, shells, jt, S2, h, 2.0, 0.4, spin, sia)
        assert orc.delta_energy(S, i, sn) == pytest.approx(ref2 - ref, rel=1e-10, abs=1e-12)
    # magnetisation: sum_i stagger * s * S_i
    assert np.allclose(orc.magnetisation(S), (spin[species - 1][:, None] * S).sum(axis=0))


def _exact_ising_4x4(beta, J=-1.0):
    L = 4
    n = L * L
    best = []
    E_list, M_list = [], []
    for state in range(1 << n):
        s = np.array([1 if (state >> k) & 1 else -1 for k in range(n)]).reshape(L, L)
        e = J * (s * np.roll(s, 1, 0) + s * np.roll(s, 1, 1)).sum()   # physical energy
        E_list.append(e)
        M_list.append(abs(s.sum()))
    E = np.array(E_list, float)
    M = np.array(M_list, float)
    w = np.exp(-beta * (E - E.min()))
    Z = w.sum()
    # reference estimators: e = E_tot/(2N) = E_phys/N (no Zeeman), m = |M|/N
    e = E / n
    m = M / n
    return (w * e).sum() / Z, (w * e * e).sum() / Z, (w * m).sum() / Z


def test_oracle_metropolis_matches_exact_enumeration():
    beta = 1.0 / 2.5
    orc, _, _ = sc_oracle(4, -1.0, model=0, Lz=1)
    n = 16
    S = np.zeros((n, 3))
    S[:, 2] = 1.0
    order = np.arange(n, dtype=np.int32)
    rng_id = np.arange(n, dtype=np.int32)
    es, ms = [], []
    for step in range(40000):
        orc.sweep(S, None, order, rng_id, beta, 99, 0, step)
        if step >= 1000:
            es.append(orc.total_energy(S)[0] / (2 * n))
            ms.append(abs(S[:, 2].sum()) / n)
    es, ms = np.array(es), np.array(ms)
    e_ex, e2_ex, m_ex = _exact_ising_4x4(beta)

    def err(x, nb=40):
        b = x[: len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
        return b.std(ddof=1) / math.sqrt(nb)

    assert abs(es.mean() - e_ex) < 4 * err(es)
    assert abs(ms.mean() - m_ex) < 4 * err(ms)
    assert abs((es ** 2).mean() - e2_ex) < 4 * err(es ** 2)


def test_single_spin_in_field_langevin():
    beta, H = 1.0, 1.7
    orc = o.Oracle(1, 1, 1, [1], [1.0], [0], [0], [0.0, 0.0, 0.0], [1.0], zeeman=[0, 0, H],
                   g_factor=1.0, mu_b=1.0)
    S = np.array([[0, 0, 1.0]])
    z = []
    for step in range(60000):
        orc.sweep(S, None, [0], [0], beta, 3, 0, step)
        z.append(S[0, 2])
    z = np.array(z[1000:])
    x = beta * H
    lang = 1 / math.tanh(x) - 1 / x
    b = z[: len(z) // 50 * 50].reshape(50, -1).mean(axis=1)
    assert abs(z.mean() - lang) < 4 * b.std(ddof=1) / math.sqrt(50)


def test_single_spin_uniaxial_anisotropy():
    from scipy.integrate import quad
    beta, D = 1.0, -2.0     # easy axis
    orc = o.Oracle(1, 1, 1, [1], [1.0], [0], [0], [0.0, 0.0, 0.0], [1.0], sia=[0, 0, D])
    S = np.array([[1.0, 0, 0]])
    z2 = []
    for step in range(60000):
        orc.sweep(S, None, [0], [0], beta, 4, 0, step)
        z2.append(S[0, 2] ** 2)
    z2 = np.array(z2[1000:])
    num = quad(lambda t: t * t * math.exp(-beta * D * t * t), -1, 1)[0]
    den = quad(lambda t: math.exp(-beta * D * t * t), -1, 1)[0]
    b = z2[: len(z2) // 50 * 50].reshape(50, -1).mean(axis=1)
    assert abs(z2.mean() - num / den) < 4 * b.std(ddof=1) / math.sqrt(50)


def test_overrelaxation_invariants():
    rng = np.random.default_rng(11)
    orc, _, _ = sc_oracle(4, [-1.0, 0.3], shells=(1.0, 2 ** 0.5))
    S = random_spins(rng, 64)
    for i in (0, 13, 63):
        A = orc.exchange_field(S, i)
        e0 = orc.site_energy(S, i)
        S1 = S.copy()
        orc.reflect(S1, i)
        assert abs(np.linalg.norm(S1[i]) - 1) < 1e-14
        assert np.dot(S1[i], A) == pytest.approx(np.dot(S[i], A), rel=1e-12, abs=1e-14)
        assert orc.site_energy(S1, i) == pytest.approx(e0, rel=1e-12, abs=1e-14)
        orc.reflect(S1, i)
        assert np.allclose(S1[i], S[i], atol=1e-13)


def test_estimators_closed_forms():
    # constant samples: C_v = chi = 0, U = 2/3 (src/get_eng.f90:50-54, src/get_mag.f90:50-53)
    e, m, n = -0.7, 0.4, 200
    mom = [n, n * e, n * e * e, n * e ** 4, n * m, n * m * m, n * m ** 4]
    out = o.repeat_observables(mom, 2.0, 100)
    assert np.allclose(out, [e, 2 / 3, 0.0, m, 2 / 3, 0.0], atol=1e-12)
    # two-value series e in {a, b}: cv = beta^2 N (a-b)^2/4
    a, b = -1.0, -0.5
    mom = [2, a + b, a * a + b * b, a ** 4 + b ** 4, 0.1 + 0.3, 0.01 + 0.09, 0.1 ** 4 + 0.3 ** 4]
    out = o.repeat_observables(mom, 3.0, 64)
    assert out[2] == pytest.approx(9.0 * 64 * (a - b) ** 2 / 4)
    assert out[5] == pytest.approx(3.0 * 64 * (0.3 - 0.1) ** 2 / 4)
    # the driver's host estimators are the same formulas
    assert np.allclose(repeat_observables(np.array(mom + [0.0]), 3.0, 64), out)
    per = np.random.default_rng(1).normal(size=(10, 6))
    mean, sem = o.over_repeats(per)
    m2, s2 = over_repeats(per)
    assert np.allclose(mean, per.mean(0)) and np.allclose(mean, m2)
    assert np.allclose(sem, per.std(0, ddof=1) / math.sqrt(10)) and np.allclose(sem, s2)


def test_pt_rule_and_two_level_acceptance():
    # src/parallel_tempering.f90:77-89: delta = -(beta_p - beta)(E - E_p)
    assert o.pt_accept(1.0, 2.0, -3.0, -1.0, 0.999)           # delta = +2 > 0
    assert not o.pt_accept(1.0, 2.0, -1.0, -3.0, 0.2)          # delta = -2, u > e^-2
    assert o.pt_accept(1.0, 2.0, -1.0, -3.0, 0.1)              # u < e^-2 = 0.135
    # two-level systems E in {0, 1} at beta1 < beta2: <acc> = E[min(1, exp(dB dE))]
    rng = np.random.default_rng(2)
    b1, b2 = 0.5, 2.0
    p1 = math.exp(-b1) / (1 + math.exp(-b1))
    p2 = math.exp(-b2) / (1 + math.exp(-b2))
    exact = 0.0
    for e1, pe1 in ((0, 1 - p1), (1, p1)):
        for e2, pe2 in ((0, 1 - p2), (1, p2)):
            exact += pe1 * pe2 * min(1.0, math.exp((b1 - b2) * (e1 - e2)))
    n = 40000
    e1 = (rng.random(n) < p1).astype(float)
    e2 = (rng.random(n) < p2).astype(float)
    u = rng.random(n)
    acc = np.mean([o.pt_accept(b1, b2, a, b, uu) for a, b, uu in zip(e1, e2, u)])
    assert abs(acc - exact) < 4 * math.sqrt(exact * (1 - exact) / n)


def test_mlp_oracle_gradient_finite_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.3, size=o.NPARAM)
    x = rng.uniform(-1, 1, 12)
    y = rng.normal(size=(12, 2))
    w = rng.uniform(0.5, 1.5, (12, 2))
    loss, g = o.mlp_loss_grad(p, x, y, w)
    for k in list(rng.choice(o.NPARAM, 40, replace=False)) + [0, 64, 128, o.NPARAM - 1]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (o.mlp_loss_grad(pp, x, y, w)[0] - o.mlp_loss_grad(pm, x, y, w)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-6, abs=1e-9)
    # forward agrees with the loss: zero loss at y = forward(x)
    yf = o.mlp_forward(p, x)
    assert o.mlp_loss_grad(p, x, yf, w)[0] == pytest.approx(0.0, abs=1e-20)


def test_structure_factor_closed_forms():
    L = 4
    pos = np.array([[i, j, k] for k in range(L) for j in range(L) for i in range(L)], float)
    S = np.tile([0.0, 0.0, 1.0], (L ** 3, 1))
    assert o.structure_factor(pos, S, [0, 0, 0]) == pytest.approx(1.0)
    assert o.structure_factor(pos, S, [2 * np.pi / L, 0, 0]) == pytest.approx(0.0, abs=1e-24)
    sign = np.where(pos.sum(axis=1) % 2 == 0, 1.0, -1.0)
    S = np.zeros_like(S)
    S[:, 0] = sign
    assert o.structure_factor(pos, S, [np.pi, np.pi, np.pi]) == pytest.approx(1.0)
    assert o.structure_factor(pos, S, [0, 0, 0]) == pytest.approx(0.0, abs=1e-24)


def test_structure_factor_parts_closed_forms():
    """The reference's whole output line (structure_factor.f90:175-177):
    an FM state along x gives {1, 1, 0, 0} at q = 0; a planar spiral
    S = (cos Q.r, sin Q.r, 0) gives |S_x|/N = |S_y|/N = 1/2 and total 1/2 at
    q = Q, and total 0 at q = 2Q; the first entry is o_structure_factor."""
    L = 8
    pos = np.array([[i, j, k] for k in range(4) for j in range(4) for i in range(L)], float)
    n = len(pos)
    S = np.tile([1.0, 0.0, 0.0], (n, 1))
    assert np.allclose(o.structure_factor_parts(pos, S, [0, 0, 0]), [1, 1, 0, 0], atol=1e-14)
    Q = np.array([2 * np.pi * 3 / L, 0, 0])
    ph = pos @ Q
    S = np.stack([np.cos(ph), np.sin(ph), np.zeros(n)], axis=1)
    got = o.structure_factor_parts(pos, S, Q)
    assert np.allclose(got, [0.5, 0.5, 0.5, 0.0], atol=1e-12)
    assert np.allclose(o.structure_factor_parts(pos, S, 2 * Q), 0.0, atol=1e-12)
    rng = np.random.default_rng(3)
    S = random_spins(rng, n)
    q = rng.normal(size=3)
    got = o.structure_factor_parts(pos, S, q)
    assert got[0] == pytest.approx(o.structure_factor(pos, S, q), rel=1e-13)
    assert got[0] == pytest.approx(float((got[1:] ** 2).sum()), rel=1e-12)


def test_ising_bond_sums_against_pair_count_from_positions():
    """o_ising_bond_sums against a count written from positions: two species,
    four shells, classes = (shell, species pair); every ordered pair (i, j),
    j != i, at a shell distance under some image adds sigma_i sigma_j to its
    class.  A wrong class index, a transposed species pair or sigma_i sigma_i
    fails it."""
    rng = np.random.default_rng(17)
    L = 3
    abc = np.eye(3)
    basis = [[0, 0, 0], [0.5, 0.5, 0.5]]
    shells = [1.0, 2 ** 0.5, 3 ** 0.5, 0.75 ** 0.5]
    cnt, idx = build_neighbours(abc, basis, (L, L, L), (0, 0, 0), shells)
    n = 2 * L ** 3
    species = np.tile([1, 2], L ** 3).astype(np.int32)
    pos = np.array([[i, j, k] for k in range(L) for j in range(L) for i in range(L)], float)
    pos = (pos[:, None, :] + np.array(basis)[None]).reshape(-1, 3)
    ns = len(shells)
    cls_of = np.arange(ns * 4).reshape(ns, 2, 2)       # class of (shell, sp_i, sp_j)
    jt = np.zeros((ns, 2, 2, 3))
    jt[..., 2] = rng.normal(size=(ns, 2, 2))
    orc = o.Oracle(0, n, 2, species, np.ones(n), cnt, idx, jt.reshape(-1), [1.0, 1.0])
    S = np.zeros((n, 3))
    S[:, 2] = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    coe = np.full(idx.shape, -1, np.int32)
    for s in range(ns):
        for i in range(n):
            for k in range(cnt[s, i]):
                coe[s, i, k] = cls_of[s, species[i] - 1, species[idx[s, i, k]] - 1]
    got = orc.ising_bond_sums(S, coe, ns * 4)
    want = np.zeros(ns * 4, np.int64)
    shifts = [np.array(v) for v in itertools.product((0, 1, -1), repeat=3)]
    for i in range(n):
        for j in range(n):
            if j == i:
                continue
            for s, d in enumerate(shells):
                if any(abs(d - np.linalg.norm(pos[i] - (pos[j] + sh * L))) <= 1e-5 for sh in shifts):
                    want[cls_of[s, species[i] - 1, species[j] - 1]] += int(S[i, 2] * S[j, 2])
    assert np.array_equal(got, want)
    # and the classes carry the whole exchange energy: sum_k J_k B_k = E_exch
    jk = jt[..., 2].reshape(-1)
    assert float(np.dot(jk, want)) == pytest.approx(orc.total_energy(S)[0], rel=1e-12, abs=1e-12)


def test_overrelax_pass_selection_rate():
    """A pass with frac = 0.3 reflects each site with probability 0.3: over
    5120 site draws the reflected count lies within 4 sigma of the binomial
    mean (an inverted compare gives 0.7, a dropped test 1.0).  Unselected
    sites keep their spin; selected ones are the reflection about the field
    of the spins present when the site is visited."""
    rng = np.random.default_rng(19)
    L = 8
    orc, _, _ = sc_oracle(L, [-1.0, 0.3], shells=(1.0, 2 ** 0.5))
    n = L ** 3
    order = np.arange(n, dtype=np.int32)
    rid = np.arange(n, dtype=np.int32)
    frac, total, hits = 0.3, 0, 0
    for step in range(10):
        S = random_spins(rng, n)
        phi = np.arctan2(S[:, 1], S[:, 0])
        S1, phi1 = S.copy(), phi.copy()
        orc.overrelax_pass(S1, phi1, order, rid, frac, 1992, 0, step, 0)
        # replay site by site: a site either keeps its spin or is reflected
        cur = S.copy()
        for i in range(n):
            if np.array_equal(S1[i], S[i]):
                continue
            tmp = cur.copy()
            orc.reflect(tmp, i)
            assert np.allclose(S1[i], tmp[i], atol=1e-13)
            cur[i] = S1[i]
            hits += 1
        total += n
    sd = math.sqrt(total * frac * (1 - frac))
    assert abs(hits - frac * total) < 4 * sd, (hits, total)


def test_start_state_oracle_and_library():
    """generate_supercell's start state (src/generate_supercell.f90:34-63):
    the library's ether_start_state equals the oracle's restatement; Ising is
    (-1)^i by file index; Marsaglia points are uniform on S^2 (<S> = 0,
    <S_z^2> = 1/3, flat S_z histogram); ANGLE has S_z ~ U(-1,1) and
    phi ~ U(0, 2 pi) with S_x = sqrt(1-S_z^2) cos phi."""
    from ether_amd.engine import start_state
    idx = np.arange(1, 6)
    S, phi = start_state(7, 0, 0, idx)
    assert np.array_equal(S[:, 2], (-1.0) ** idx) and not S[:, :2].any()
    nb = 20000
    fi = np.arange(1, nb + 1)
    for angle in (0, 1):
        S, phi = start_state(1992, 1, angle, fi)
        So, po = o.start_state(1992, 1, angle, fi)
        assert np.allclose(S, So, atol=1e-15, rtol=0) and np.allclose(phi, po, atol=1e-15, rtol=0)
        assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-14)
        sd = 1.0 / math.sqrt(3 * nb)
        assert np.all(np.abs(S.mean(axis=0)) < 4.5 * sd)
        assert abs((S[:, 2] ** 2).mean() - 1 / 3) < 4.5 * math.sqrt(4 / 45 / nb)
        h, _ = np.histogram(S[:, 2], bins=20, range=(-1, 1))
        assert ((h - nb / 20) ** 2 / (nb / 20)).sum() < 45.0      # chi^2, 19 dof, p ~ 1e-3
        if angle:
            assert np.all((phi >= 0) & (phi < 2 * math.pi))
            assert np.allclose(S[:, 0], np.sqrt(1 - S[:, 2] ** 2) * np.cos(phi), atol=1e-12)
            hp, _ = np.histogram(phi, bins=20, range=(0, 2 * math.pi))
            assert ((hp - nb / 20) ** 2 / (nb / 20)).sum() < 45.0
        else:
            assert not phi.any()
    # a function of the seed
    assert not np.allclose(start_state(1, 1, 0, fi[:8])[0], start_state(2, 1, 0, fi[:8])[0])
