"""Batches of uneven lattices: one launch per kernel kind for all of them.

Each lattice is an ordinary engine (its own sites, couplings, replicas,
temperatures, ladders and seed).  A Batch advances all of them through the
schedule of ether_run with batched launches and graph replays; every
lattice's spins, moments and tempering state must be bit-identical to the
same engine run on its own (spiral-state sums included where a lattice has
spiral partners).
"""
import numpy as np
import pytest

from ether_amd import engine as E
from ether_amd import workloads
from ether_amd.driver import attach_neighbours, make_engine
from tests.helpers import random_xyz, say, three_shell_ising

pytestmark = pytest.mark.gpu


def _engine(name, R, seed, over, temps):
    if name == "sc_3shell":
        m = three_shell_ising(over["L"])
    else:
        m = workloads.model(name, **over)
        attach_neighbours(m)
    eng = make_engine(m, R, seed=seed)
    eng.set_betas(m.betas(temps))
    eng.set_ladders([0, R])
    rng = np.random.default_rng(seed)
    for r in range(R):
        if m.tags.ising:
            S = np.zeros((m.n_sites, 3))
            S[:, 2] = np.where(rng.random(m.n_sites) < 0.5, 1.0, -1.0)
            eng.set_spins(r, S)
        else:
            S = random_xyz(rng, m.n_sites)
            eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))
    return m, eng


def _state(eng):
    spins = np.stack([eng.get_spins(r)[0] for r in range(eng.R)])
    extra = [np.asarray(x) for x in eng.pt_state()]
    if getattr(eng, "spiral_count", 0):
        extra += list(eng.spiral())
    return spins, eng.moments(), extra


XYZ = [
    ("sc_heis_afm_32", 20, 11, dict(sc="6 6 6", ssp=".T.", ss_dis="1.01"), np.linspace(60, 6, 20)),
    ("tri_afm_96", 13, 12, dict(sc="12 12 1", sia=".F.", ssp=".T.", ss_dis="1.01", ss_direc="1 0 0"),
     np.linspace(20, 2, 13)),                                                              # ANGLE, 3 colours
    ("pyro_pt_16", 16, 13, dict(sc="2 2 2", pt=".T. 11"), np.geomspace(20, 1, 16)),        # 4 colours, D 18
    ("sc_heis_afm_32", 70, 14, dict(sc="4 4 4"), np.linspace(40, 4, 70)),                  # two replica groups
]
ISING = [
    ("sc_ising_16", 30, 21, dict(sc="6 6 6"), np.linspace(100, 10, 30)),
    ("pyro_pt_16", 20, 22, dict(sc="2 2 2", model="ising", para=".T. 1", pt=".T. 11",
                                zeeman=".T. 0 0 0.3"), np.linspace(6, 2, 20)),
]


@pytest.mark.parametrize("cases,graphs", [(XYZ, True), (XYZ, False), (ISING, True)])
def test_batch_equals_lattices_run_alone(cases, graphs):
    say(f"batch of {len(cases)} lattices, graphs={graphs}")
    alone, batched = [], []
    for name, R, seed, over, temps in cases:
        m, a = _engine(name, R, seed, over, temps)
        _, b = _engine(name, R, seed, over, temps)
        alone.append((m, a))
        batched.append(b)
    batch = E.Batch(batched)
    batch.set_graphs(graphs)
    # two calls: 37 MCS (a partial period at the end) and 63 MCS; OVRR every MCS
    # (fused), samples every 5 MCS after 20, swaps every 5 MCS
    sc = 0
    sc = batch.run(1, 37, 20, 5, 1, 1.0, 5, sc)
    sc = batch.run(38, 63, 20, 5, 1, 1.0, 5, sc)
    for m, a in alone:
        sca = a.run(1, 37, 20, 5, 1, 1.0, 5, 0)
        sca = a.run(38, 63, 20, 5, 1, 1.0, 5, sca)
        assert sca == sc
    for (m, a), b in zip(alone, batched):
        sa, ma, pa = _state(a)
        sb, mb, pb = _state(b)
        assert np.array_equal(sa, sb), f"{m.tags.model} N={m.n_sites}: spins differ"
        assert np.array_equal(ma, mb), f"{m.tags.model} N={m.n_sites}: moments differ"
        for x, y in zip(pa, pb):
            assert np.array_equal(x, y), f"{m.tags.model} N={m.n_sites}: tempering/spiral state differs"
        assert ma[:, E.MOM_N].sum() > 0
    batch.close()
    for (m, a), b in zip(alone, batched):
        a.close()
        b.close()


# Ising batches sweep with every replica's spins resident in LDS (a
# workgroup per replica of every lattice, the sweeps between two measurements
# in one launch).  Cases: 40^3 (64,000 sites: above the one-context resident
# budget, so alone it runs the phase kernels), an odd edge (more than two
# colours), two coupling classes with Zeeman, three classes (26 neighbours:
# the kernel instance for any number of classes, launched beside the others),
# and 6^3.
ISING_RES = [
    ("sc_ising_16", 8, 31, dict(sc="40 40 40"), np.linspace(100, 10, 8)),
    ("sc_ising_16", 12, 32, dict(sc="9 9 9"), np.linspace(100, 10, 12)),
    ("pyro_pt_16", 20, 22, dict(sc="2 2 2", model="ising", para=".T. 1", pt=".T. 11",
                                zeeman=".T. 0 0 0.3"), np.linspace(6, 2, 20)),
    ("sc_3shell", 12, 23, dict(L=6), np.linspace(6, 2, 12)),
    ("sc_ising_16", 30, 21, dict(sc="6 6 6"), np.linspace(100, 10, 30)),
]


@pytest.mark.parametrize("graphs", [True, False])
def test_resident_batch_equals_phase_batch_and_alone(graphs):
    say(f"resident Ising batch of {len(ISING_RES)} lattices, graphs={graphs}")
    res, phase, alone = [], [], []
    for name, R, seed, over, temps in ISING_RES:
        m, a = _engine(name, R, seed, over, temps)
        res.append(a)
        _, b = _engine(name, R, seed, over, temps)
        b.set_resident(0)
        phase.append(b)
        _, c = _engine(name, R, seed, over, temps)
        alone.append(c)
    batches = [E.Batch(res), E.Batch(phase)]
    scs = []
    for batch in batches:
        batch.set_graphs(graphs)
        sc = batch.run(1, 37, 20, 5, 0, 0.0, 5, 0)
        sc = batch.run(38, 63, 20, 5, 0, 0.0, 5, sc)
        scs.append(sc)
    assert batches[0].resident() and not batches[1].resident()
    sca = 0
    for c in alone:
        sca = c.run(1, 37, 20, 5, 0, 0.0, 5, 0)
        sca = c.run(38, 63, 20, 5, 0, 0.0, 5, sca)
    assert scs[0] == scs[1] == sca
    for k, (name, R, *_rest) in enumerate(ISING_RES):
        sa, ma, pa = _state(res[k])
        for other in (phase[k], alone[k]):
            sb, mb, pb = _state(other)
            assert np.array_equal(sa, sb), f"{name} R={R}: spins differ"
            assert np.array_equal(ma, mb), f"{name} R={R}: moments differ"
            for x, y in zip(pa, pb):
                assert np.array_equal(x, y), f"{name} R={R}: tempering state differs"
        assert ma[:, E.MOM_N].sum() > 0
        # the last sweep's acceptance per replica (ether_sweep)
        assert np.array_equal(res[k].sweep(2, want_accept=True), phase[k].sweep(2, want_accept=True))
    for batch in batches:
        batch.close()
    for e in res + phase + alone:
        e.close()


def test_contexts_continue_alone_after_their_batch():
    """Two batches (Ising resident, XYZ phases) run back to back on their two
    streams, then every context continues with its own ether_run at once: the
    wait for its batch is made at that first use of the context's stream, so
    the spins, moments and tempering state equal a context that ran the whole
    schedule alone."""
    say("batches, then each context alone")
    cases = ISING_RES[1:3] + XYZ[:2]
    mixed, ref = [], []
    for name, R, seed, over, temps in cases:
        _, a = _engine(name, R, seed, over, temps)
        _, b = _engine(name, R, seed, over, temps)
        mixed.append(a)
        ref.append(b)
    bi = E.Batch([e for e in mixed if e.model == E.MODEL_ISING])
    bx = E.Batch([e for e in mixed if e.model != E.MODEL_ISING])
    sc_i = bi.run(1, 30, 10, 5, 1, 1.0, 5, 0)
    sc_x = bx.run(1, 30, 10, 5, 1, 1.0, 5, 0)
    for e in mixed:
        e.run(31, 17, 10, 5, 1, 1.0, 5, sc_i if e.model == E.MODEL_ISING else sc_x)
    for e in ref:
        sc = e.run(1, 30, 10, 5, 1, 1.0, 5, 0)
        e.run(31, 17, 10, 5, 1, 1.0, 5, sc)
    for (name, R, *_rest), a, b in zip(cases, mixed, ref):
        sa, ma, pa = _state(a)
        sb, mb, pb = _state(b)
        assert np.array_equal(sa, sb), f"{name} R={R}: spins differ"
        assert np.array_equal(ma, mb), f"{name} R={R}: moments differ"
        for x, y in zip(pa, pb):
            assert np.array_equal(x, y), f"{name} R={R}: tempering state differs"
    bi.close()
    bx.close()
    for e in mixed + ref:
        e.close()
