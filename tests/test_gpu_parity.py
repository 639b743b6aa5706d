"""GPU parity against the CPU oracle, through the C-ABI.

Ising: bit-exact.  The device and the oracle draw the same Philox words and
compare them with the same integer thresholds, so after K sweeps every spin of
every replica must be identical, and the integer bond sums behind E_tot must
be equal.

Heisenberg (FP32 on the device, FP64 in the oracle): local fields and site
energies within 1e-6 of the oracle relative to the field scale sum_j |J_j|;
energies within 1e-6 relative to sum_i |e_i|; after one sweep every site whose
Metropolis margin |-beta dE - ln eta| exceeds 1e-3 makes the same decision and
lands within 2e-6 of the oracle's spin.
"""
import math

import numpy as np
import pytest

from tests.helpers import colour_order, pair, random_xyz, say
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def _ising_case(name, R, K, **over):
    say(f"ising parity {name} {over} R={R} K={K}")
    m, eng, orc = pair(name, R, **over)
    rng = np.random.default_rng(7)
    temps = np.linspace(2.0, 6.0, R) if m.tags.para else np.linspace(10.0, 100.0, R)
    betas = m.betas(temps)
    eng.set_betas(betas)
    states = []
    for r in range(R):
        S = np.zeros((m.n_sites, 3))
        S[:, 2] = np.where(rng.random(m.n_sites) < 0.5, 1.0, -1.0)
        eng.set_spins(r, S)
        states.append(S)
    eng.sweep(K)
    nc, col, order, rid = colour_order(eng)
    for r in range(R):
        S = states[r]
        acc = 0
        for step in range(K):
            acc += orc.sweep(S, None, order, rid, betas[r], m.tags.seed, r, step)
        Sg, _ = eng.get_spins(r)
        assert np.array_equal(Sg[:, 2], S[:, 2]), f"replica {r}: {np.sum(Sg[:, 2] != S[:, 2])} spins differ"
    # energies: exact integer bond sums per coupling class; magnetisation too
    _, M = eng.measure()
    parts, bonds = eng.energy_parts()
    coe, ncls = ising_classes(m)
    assert bonds.shape[1] == ncls
    for r in range(R):
        S = states[r]
        assert np.allclose(M[r], orc.magnetisation(S), rtol=0, atol=1e-9)
        ex, ze, si = orc.total_energy(S)
        assert parts[r, 0] == pytest.approx(ex, rel=1e-12, abs=1e-12)
        assert parts[r, 1] == pytest.approx(ze, rel=1e-12, abs=1e-12)
        # every coupling class: the oracle's integer bond sums of the same classes
        assert np.array_equal(bonds[r], orc.ising_bond_sums(S, coe, ncls)), f"replica {r}"
        if bonds.shape[1] == 1:
            # one class: B = E_exch / J exactly
            jzz = m.j_exc.reshape(-1, 3)[np.flatnonzero(m.j_exc.reshape(-1, 3)[:, 2])[0], 2]
            assert bonds[r, 0] == round(ex / jzz)
    eng.close()


def ising_classes(m):
    """The engine's Ising coupling classes: distinct nonzero J_zz numbered in
    order of first appearance over sites, shells and neighbours; returns the
    class of every neighbour entry (-1 for zero couplings) and the count."""
    ns, n = m.nn_count.shape
    nsp = m.n_species
    jz = m.j_exc.reshape(ns, nsp, nsp, 3)[..., 2]
    seen = {}
    coe = np.full(m.nn_idx.shape, -1, np.int32)
    for i in range(n):
        a = m.species[i] - 1
        for s in range(ns):
            for k in range(m.nn_count[s, i]):
                J = jz[s, a, m.species[m.nn_idx[s, i, k]] - 1]
                if J != 0.0:
                    coe[s, i, k] = seen.setdefault(float(J), len(seen))
    return coe, len(seen)


def test_ising_sc_bit_exact():
    _ising_case("sc_ising_16", 8, 6, sc="8 8 8", para=".T. 1")


def test_ising_sc_mev_units_bit_exact():
    _ising_case("sc_ising_16", 4, 4, sc="6 6 6")


def test_ising_pyrochlore_two_shells_zeeman_bit_exact():
    _ising_case("pyro_pt_16", 4, 4, sc="2 2 2", model="ising", para=".T. 1", pt=".F. 11",
                zeeman=".T. 0 0 0.3")


# 128 replicas or more: the measurement pass holds four replicas per lane
def test_ising_sc_many_replicas_bit_exact():
    _ising_case("sc_ising_16", 150, 2, sc="6 6 6")


def test_ising_sc_full_size_bit_exact():
    # 16^3 sites, 190 replicas: three waves of replicas per site, four per lane
    # in the measurement
    _ising_case("sc_ising_16", 190, 2)


def test_ising_pyrochlore_many_replicas_bit_exact():
    _ising_case("pyro_pt_16", 130, 2, sc="2 2 2", model="ising", para=".T. 1", pt=".F. 11",
                zeeman=".T. 0 0 0.3")


def _xyz_state(m, eng, R, seed):
    rng = np.random.default_rng(seed)
    states = []
    for r in range(R):
        S = random_xyz(rng, m.n_sites).astype(np.float32).astype(np.float64)
        phi = np.arctan2(S[:, 1], S[:, 0]).astype(np.float32).astype(np.float64)
        eng.set_spins(r, S, phi)
        states.append((S, phi))
    return states


@pytest.mark.parametrize("name,over", [
    ("sc_heis_afm_32", dict(sc="8 8 8")),
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11")),
    ("tri_afm_96", dict(sc="12 12 1")),
])
def test_xyz_local_fields_and_energy(name, over):
    say(f"xyz fields {name} {over}")
    m, eng, orc = pair(name, 2, **over)
    eng.set_betas(m.betas([5.0, 10.0]))
    states = _xyz_state(m, eng, 2, 3)
    jabs = np.abs(m.j_exc).max() * m.nn_count.sum(axis=0).max()
    for r in range(2):
        S, _ = states[r]
        h, e = eng.local_fields(r)
        ho = np.array([orc.exchange_field(S, i) for i in range(m.n_sites)])
        eo = np.array([orc.site_energy(S, i) for i in range(m.n_sites)])
        assert np.max(np.abs(h - ho)) <= 1e-6 * jabs
        escale = jabs + np.abs(m.h).max() * m.mu_b * m.g_factor + (
            np.abs(m.sia_vec).max() if m.sia_vec is not None else 0.0)
        assert np.max(np.abs(e - eo)) <= 1e-6 * escale
    eng.measure()
    parts, _ = eng.energy_parts()
    for r in range(2):
        S, _ = states[r]
        ex, ze, si = orc.total_energy(S)
        scale = sum(abs(orc.site_energy(S, i)) for i in range(m.n_sites))
        assert abs(parts[r, 0] - ex) <= 1e-6 * scale
        assert abs(parts[r, 1] - ze) <= 1e-6 * scale + 1e-300
        assert abs(parts[r, 2] - si) <= 1e-6 * scale + 1e-300
        assert parts[r, 3] == pytest.approx(0.5 * parts[r, 0] + parts[r, 1] + parts[r, 2])
    eng.close()


def _oracle_sweep_with_margins(m, orc, states, order, rid, betas, seed, step):
    """The oracle's sweep of all replicas, site by site with the device's proposal
    rule (pooled Marsaglia candidates), recording -beta dE - ln eta."""
    R = len(states)
    S = np.stack([s for s, _ in states]).copy()
    phi = np.stack([p for _, p in states]).copy()
    marg = np.zeros((R, m.n_sites))
    for i in order:
        sn, pn, eta = orc.propose_site(seed, int(rid[i]), step, S[:, i, :], phi[:, i])
        for r in range(R):
            dE = orc.delta_energy(S[r], i, sn[r])
            marg[r, i] = -betas[r] * dE - math.log(eta[r])
            if marg[r, i] > 0:
                S[r, i] = sn[r]
                phi[r, i] = pn[r]
    return S, phi, marg


@pytest.mark.parametrize("name,over,R", [
    ("sc_heis_afm_32", dict(sc="8 8 8"), 2),
    ("sc_heis_afm_32", dict(sc="4 4 4"), 70),                 # two pooling groups, one partial
    ("tri_afm_96", dict(sc="12 12 1"), 2),                    # ANGLE + Zeeman + SIA
    ("tri_afm_96", dict(sc="12 12 1", angle=".F. 0"), 3),     # Marsaglia + Zeeman + SIA
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 2),
])
def test_xyz_one_sweep_matches_oracle(name, over, R):
    say(f"xyz sweep {name} {over} R={R}")
    m, eng, orc = pair(name, R, **over)
    lo, hi = (2.0, 8.0) if name == "pyro_pt_16" else (5.0, 20.0)
    betas = m.betas(np.linspace(lo, hi, R))
    eng.set_betas(betas)
    states = _xyz_state(m, eng, R, 5)
    eng.sweep(1)
    nc, col, order, rid = colour_order(eng)
    nbrs = [set() for _ in range(m.n_sites)]
    for s in range(m.nn_count.shape[0]):
        for i in range(m.n_sites):
            nbrs[i].update(m.nn_idx[s, i, : m.nn_count[s, i]].tolist())
    S, phi, marg = _oracle_sweep_with_margins(m, orc, states, order, rid, betas, m.tags.seed, 0)
    n_checked = 0
    for r in range(R):
        Sg, phig = eng.get_spins(r)
        tainted = set()
        for i in order:
            if i in tainted:
                continue
            dev = np.abs(Sg[i] - S[r, i]).max()
            if dev > 2e-6:
                assert abs(marg[r, i]) < 1e-3, f"replica {r} site {i}: margin {marg[r, i]}, dev {dev}"
                tainted |= nbrs[i]
            else:
                n_checked += 1
    assert n_checked > 0.99 * R * m.n_sites
    eng.close()


@pytest.mark.parametrize("frac", [1.0, 0.5, 0.3])
def test_overrelaxation_matches_oracle(frac):
    say(f"ovrr frac={frac}")
    m, eng, orc = pair("sc_heis_afm_32", 2, sc="6 6 6")
    eng.set_betas(m.betas([5.0, 10.0]))
    states = _xyz_state(m, eng, 2, 9)
    eng.measure()
    e_before = eng.energy_parts()[0][:, 3].copy()
    eng.overrelax(frac)
    eng.measure()
    e_after = eng.energy_parts()[0][:, 3]
    nc, col, order, rid = colour_order(eng)
    for r in range(2):
        S, phi = states[r]
        S = S.copy()
        phi = phi.copy()
        orc.overrelax_pass(S, phi, order, rid, frac, m.tags.seed, r, 0, 0)
        Sg, _ = eng.get_spins(r)
        assert np.max(np.abs(Sg - S)) < 2e-5
        assert np.allclose(np.linalg.norm(Sg, axis=1), 1.0, atol=1e-6)
    # reflections conserve the energy
    assert np.allclose(e_after, e_before, rtol=1e-5, atol=1e-9)
    eng.close()


def _oracle_fused_with_margins(m, orc, states, order, rid, betas, seed, step, frac):
    """The fused schedule on the oracle: site by site in colour order, each
    replica's spin is reflected about its local field (selected with
    probability frac by the OVRR draw) and then gets the pooled Metropolis
    proposal.  Sites of one colour share no bonds, so this order is the
    device's colour phase.  Records -beta dE - ln eta, |A| and the in-plane
    length of the reflected spin per site."""
    R = len(states)
    S = np.stack([s for s, _ in states]).copy()
    phi = np.stack([p for _, p in states]).copy()
    marg = np.zeros((R, m.n_sites))
    amag = np.zeros((R, m.n_sites))
    rho = np.ones((R, m.n_sites))
    for i in order:
        for r in range(R):
            amag[r, i] = np.linalg.norm(orc.exchange_field(S[r], i))
            if frac < 1.0:
                w = o.rng(seed, int(rid[i]), r, step, o.RNG_OVRR, 0)
                if not ((int(w[0]) >> 8) * 2.0 ** -24 < frac):
                    continue
            orc.reflect(S[r], i)
            phi[r, i] = math.atan2(S[r, i, 1], S[r, i, 0])
            rho[r, i] = math.hypot(S[r, i, 0], S[r, i, 1])
        sn, pn, eta = orc.propose_site(seed, int(rid[i]), step, S[:, i, :], phi[:, i])
        for r in range(R):
            dE = orc.delta_energy(S[r], i, sn[r])
            marg[r, i] = -betas[r] * dE - math.log(eta[r])
            if marg[r, i] > 0:
                S[r, i] = sn[r]
                phi[r, i] = pn[r]
    return S, phi, marg, amag, rho


@pytest.mark.parametrize("name,over,R,frac", [
    ("sc_heis_afm_32", dict(sc="8 8 8"), 2, 1.0),
    ("sc_heis_afm_32", dict(sc="4 4 4"), 70, 0.5),                      # OVRR draw, two groups
    ("sc_heis_afm_32", dict(sc="6 6 6"), 3, 0.3),                       # selection below one half
    ("tri_afm_96", dict(sc="12 12 1", sia=".F."), 2, 1.0),              # ANGLE + Zeeman
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 2, 1.0),
])
def test_fused_overrelaxation_sweep_matches_oracle(name, over, R, frac):
    """One MCS of ether_run with OVRR due: the reflection is fused into each
    colour phase of the sweep.  Every site whose Metropolis margin exceeds 1e-3
    must land within the FP32 error of the reflection of the oracle's spin.
    That error grows as 1/|A|; with ANGLE moves the proposal starts from
    phi = atan2(S_y, S_x) of the reflected spin, whose error grows as 1/rho
    with rho the reflected spin's in-plane length."""
    say(f"fused ovrr+sweep {name} {over} R={R} frac={frac}")
    m, eng, orc = pair(name, R, **over)
    lo, hi = (2.0, 8.0) if name == "pyro_pt_16" else (5.0, 20.0)
    betas = m.betas(np.linspace(lo, hi, R))
    eng.set_betas(betas)
    states = _xyz_state(m, eng, R, 11)
    eng.run(1, 1, 10 ** 6, 10 ** 6, 1, frac)
    nc, col, order, rid = colour_order(eng)
    nbrs = [set() for _ in range(m.n_sites)]
    for s in range(m.nn_count.shape[0]):
        for i in range(m.n_sites):
            nbrs[i].update(m.nn_idx[s, i, : m.nn_count[s, i]].tolist())
    S, phi, marg, amag, rho = _oracle_fused_with_margins(m, orc, states, order, rid, betas,
                                                         m.tags.seed, 0, frac)
    jabs = np.abs(m.j_exc).max() * m.nn_count.sum(axis=0).max()
    n_checked = 0
    for r in range(R):
        Sg, _ = eng.get_spins(r)
        tainted = set()
        for i in order:
            if i in tainted:
                continue
            tol = 2e-6 * (1.0 + jabs / max(amag[r, i], 1e-3))
            if m.tags.angle:
                tol *= 1.0 + 1.0 / max(rho[r, i], 1e-3)
            dev = np.abs(Sg[i] - S[r, i]).max()
            if dev > tol:
                assert abs(marg[r, i]) < 1e-3, f"replica {r} site {i}: margin {marg[r, i]}, dev {dev}"
                tainted |= nbrs[i]
            else:
                n_checked += 1
    assert n_checked > 0.99 * R * m.n_sites
    eng.close()


def test_split_ovrr_compat_equals_separate_calls():
    """ETHER_COMPAT_SPLIT_OVRR: ether_run performs the reference's order, a whole
    OVRR pass and then the sweep, bit for bit what ether_overrelax followed by
    ether_sweep gives."""
    from ether_amd import engine as E
    from ether_amd.driver import make_engine
    say("split ovrr compat")
    m, eng, _ = pair("sc_heis_afm_32", 3, sc="6 6 6")
    ref = make_engine(m, 3, compat=E.COMPAT_SPLIT_OVRR)
    betas = m.betas([5.0, 10.0, 20.0])
    for e in (eng, ref):
        e.set_betas(betas)
    st = _xyz_state(m, eng, 3, 21)
    for r, (S, phi) in enumerate(st):
        ref.set_spins(r, S, phi)
    eng.overrelax(1.0)
    eng.sweep(1)
    ref.run(1, 1, 10 ** 6, 10 ** 6, 1, 1.0)
    for r in range(3):
        a, _ = eng.get_spins(r)
        b, _ = ref.get_spins(r)
        assert np.array_equal(a, b)
    eng.close()
    ref.close()


@pytest.mark.parametrize("etot", [False, True])
def test_pt_swaps_match_oracle(etot):
    """Swap decisions of 24 rounds against o_pt_accept (src/parallel_tempering.f90:77-89)
    on the physical energy, or with ETHER_COMPAT_PT_ETOT on the reference's
    E_tot (exchange double counted, Zeeman once: src/ether.f90:177,194)."""
    from ether_amd import engine as EE
    from ether_amd.driver import make_engine
    say(f"pt swaps etot={etot}")
    R = 8
    over = dict(sc="4 4 4", para=".T. 1")
    if etot:
        over["zeeman"] = ".T. 0 0 0.8"
    m, eng, orc = pair("sc_heis_afm_32", R, **over)
    if etot:
        eng.close()
        eng = make_engine(m, R, seed=1992, compat=EE.COMPAT_PT_ETOT)
    betas = np.linspace(0.1, 0.45, R)
    eng.set_betas(betas)
    eng.set_ladders([0, R])
    states = _xyz_state(m, eng, R, 13)
    E = []
    for S, _ in states:
        ex, ze, si = orc.total_energy(S)
        E.append(ex + ze + si if etot else 0.5 * ex + ze + si)
    E = np.array(E)
    rep_at = list(range(R))
    total = 0
    for call in range(24):
        parity = call % 2
        n_acc = eng.pt_swap(parity)
        acc = 0
        for k in range(parity, R - 1, 2):
            ra, rb = rep_at[k], rep_at[k + 1]
            u = (int(o.rng(m.tags.seed, k, 0, call, o.RNG_PT)[0]) >> 8) / 2.0 ** 24
            if o.pt_accept(betas[k], betas[k + 1], E[ra], E[rb], u):
                rep_at[k], rep_at[k + 1] = rb, ra
                acc += 1
        assert n_acc == acc, f"call {call}"
        total += acc
        slot_of, pacc, patt, trips = eng.pt_state()
        expect = np.empty(R, np.int32)
        for k, r in enumerate(rep_at):
            expect[r] = k
        assert np.array_equal(slot_of, expect), f"call {call}"
    assert 0 < total < 24 * 4
    eng.close()


def test_pyrochlore_at_full_size_with_swaps():
    """16^3 pyrochlore cells (65536 spins, four colours), four ladders of 64
    replicas, OVRR every MCS and swaps every 11 MCS over 22 MCS: each
    ladder's slots stay a permutation of its own range, swaps were tried in
    every ladder, spins stay unit vectors, and the measurement pass's
    exchange energy matches the oracle's FP64 sum on three replicas."""
    say("pyrochlore 16^3 x 16, 4 x 64 replicas, swaps every 11 MCS")
    R, L = 256, 64
    m, eng, orc = pair("pyro_pt_16", R)
    N = m.n_sites
    assert N == 16 ** 3 * 16
    eng.set_betas(m.betas(np.tile(np.geomspace(1.0, 0.05, L), R // L)))
    eng.set_ladders(list(range(0, R + 1, L)))
    rng = np.random.default_rng(23)
    for r in range(R):
        S = random_xyz(rng, N)
        eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))
    eng.run(1, 22, 1000, 1000, 1, 1.0, 11)
    slot, acc, att, _ = eng.pt_state()
    for k in range(R // L):
        assert np.array_equal(np.sort(slot[k * L:(k + 1) * L]), np.arange(k * L, (k + 1) * L))
        assert att[k * L:(k + 1) * L].sum() > 0
    assert np.all(acc <= att)
    eng.measure()
    parts, _ = eng.energy_parts()
    for r in (0, 131, R - 1):
        S, _ = eng.get_spins(r)
        assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-5)
        ex, _, _ = orc.total_energy(S)
        scale = sum(abs(orc.site_energy(S, i)) for i in range(N))
        assert abs(parts[r, 0] - ex) <= 1e-6 * scale, (r, parts[r, 0], ex)
    eng.close()


def test_ising_refuses_more_than_255_bonds_of_one_class():
    """The Ising kernels count a site's down neighbours per coupling class in
    one byte per replica: a star of 300 sites (site 0 bonded to all others with
    one J) is refused at commit with ETHER_ERR_LIMIT instead of overflowing;
    a star of 200 commits."""
    from ether_amd import engine as EE
    say("ising class overflow refusal")
    for n, ok in ((300, False), (200, True)):
        cap = n - 1
        cnt = np.zeros((1, n), np.int32)
        idx = np.full((1, n, cap), -1, np.int32)
        cnt[0, 0] = n - 1
        idx[0, 0, : n - 1] = np.arange(1, n)
        cnt[0, 1:] = 1
        idx[0, 1:, 0] = 0
        eng = EE.Engine(EE.MODEL_ISING, n, 4, 1)
        try:
            if ok:
                eng.setup(1, np.ones(n, np.int32), np.ones(n), cnt, idx, [0, 0, 1.0], [1.0])
            else:
                with pytest.raises(EE.EtherError) as ei:
                    eng.setup(1, np.ones(n, np.int32), np.ones(n), cnt, idx, [0, 0, 1.0], [1.0])
                assert ei.value.code == EE.ETHER_ERR_LIMIT and "255" in str(ei.value)
        finally:
            eng.close()
