"""The batched uneven workload at its full size: the committed list of 256
lattices (tests/golden/c5_lattices.json; simple-cubic Ising and Heisenberg
8..48, triangular 24..128, pyrochlore 4..12 cells, random-sign J, ladders of
8..32 replicas) as two batches (Ising, Heisenberg) for 25 MCS with swaps
every 5 MCS and samples every 10 MCS after 10, graph replays for the whole
periods and launches one by one for the rest.

  - every lattice's slots stay a permutation of its ladder, swaps were tried
    on every ladder, accepted <= attempted;
  - spins stay unit vectors (XYZ) or +-1 (Ising) on the first and last
    replica of every lattice;
  - eight lattices (two of each kind) also run alone through ether_run:
    spins, moments and tempering state bit-identical to the batch;
  - on those eight, the measured energy of one replica against the oracle's
    FP64 sum (Ising: exact; XYZ: 1e-6 of sum_i |e_i|).
"""
import json
import os

import numpy as np
import pytest

from ether_amd import engine as E
from ether_amd import workloads
from ether_amd.driver import attach_neighbours, make_engine
from tests.helpers import say
from oracle import oracle as o

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _engine(sp):
    m = workloads.c5_model(sp)
    attach_neighbours(m)
    eng = make_engine(m, sp["R"], seed=sp["seed"])
    eng.set_betas(m.betas(workloads.c5_temperatures(sp)))
    eng.set_ladders([0, sp["R"]])
    eng.init_spins(E.INIT_RANDOM)
    return m, eng


def test_c5_workload_as_batches():
    with open(os.path.join(HERE, "golden", "c5_lattices.json")) as f:
        specs = json.load(f)["lattices"]
    assert len(specs) == 256
    # two lattices of each kind also run alone
    twins = []
    for kind in workloads.C5_KINDS:
        twins += [k for k, s in enumerate(specs) if s["kind"] == kind][:2]
    say(f"c5: building 256 lattices; alone: {twins}")
    models, engines, alone = [], [], {}
    for k, sp in enumerate(specs):
        m, eng = _engine(sp)
        models.append(m)
        engines.append(eng)
        if k in twins:
            alone[k] = _engine(sp)[1]
    groups = {"ising": [k for k, m in enumerate(models) if m.tags.ising],
              "xyz": [k for k, m in enumerate(models) if not m.tags.ising]}
    say(f"c5: batches of {len(groups['ising'])} Ising and {len(groups['xyz'])} Heisenberg lattices")
    batches = [E.Batch([engines[k] for k in idx]) for idx in groups.values()]
    sched = (10, 10, 0, 0.0, 5)          # mcs_eq, coa, OVRR interval, OVRR passes, swap interval
    sc = 0
    for b in batches:
        s1 = b.run(1, 25, *sched, 0)
        sc = s1
    for k, e in alone.items():
        assert e.run(1, 25, *sched, 0) == sc
    say("c5: checking every lattice")
    for k, (m, e) in enumerate(zip(models, engines)):
        R = specs[k]["R"]
        slot, acc, att, _ = e.pt_state()
        assert np.array_equal(np.sort(slot), np.arange(R)), k
        assert att.sum() > 0 and np.all(acc <= att), k
        for r in (0, R - 1):
            S, _ = e.get_spins(r)
            if m.tags.ising:
                assert np.all(np.abs(S[:, 2]) == 1.0) and not S[:, :2].any(), k
            else:
                assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-5), k
    say("c5: lattices alone against the batch and the oracle")
    for k, a in alone.items():
        b = engines[k]
        m = models[k]
        for r in range(a.R):
            assert np.array_equal(a.get_spins(r)[0], b.get_spins(r)[0]), (k, r)
        assert np.array_equal(a.moments(), b.moments()), k
        for x, y in zip(a.pt_state(), b.pt_state()):
            assert np.array_equal(x, y), k
        a.measure()
        parts, _ = a.energy_parts()
        S, _ = a.get_spins(0)
        orc = o.Oracle.from_model(m)
        ex, ze, si = orc.total_energy(S)
        if m.tags.ising:
            assert parts[0, 0] == pytest.approx(ex, rel=1e-12, abs=1e-12), k
        else:
            scale = sum(abs(orc.site_energy(S, i)) for i in range(m.n_sites))
            assert abs(parts[0, 0] - ex) <= 1e-6 * scale, (k, parts[0, 0], ex)
    for b in batches:
        b.close()
    for e in engines + list(alone.values()):
        e.close()
