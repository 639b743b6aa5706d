"""Captured HIP graphs of the MCS schedule against eager launches.

The same run (OVRR, sweeps, tempering swaps, samples; equilibration and
production periods and an unaligned tail) is executed once with graphs and
once launch by launch, from the same state.  Every result must be identical
(spins, moments, slot permutation, swap counters), also across several
replays of the same graph with host work and synchronisations in between.
"""
import numpy as np
import pytest

from ether_amd import engine as E
from tests.helpers import pair, say

pytestmark = pytest.mark.gpu


def _run(name, graphs, chunks, **over):
    m, eng, _ = pair(name, 12, **over)
    eng.set_graphs(graphs)
    temps = np.linspace(2.0, 12.0, 12) if not m.tags.para else np.linspace(0.5, 3.0, 12)
    eng.set_betas(m.betas(temps))
    eng.set_ladders([0, 6, 12])
    eng.init_spins(E.INIT_RANDOM)
    eng.reset_moments()
    sc = 0
    step = 1
    for n in chunks:
        sc = eng.run(step, n, 100, 10, 1 if m.tags.xyz else 0, 0.6 if m.tags.xyz else 0.0, 11, sc)
        eng.synchronize()             # host work between replays
        step += n
    spins = [eng.get_spins(r)[0] for r in range(12)]
    out = dict(spins=np.array(spins), mom=eng.moments(), pt=eng.pt_state(), sc=sc)
    eng.close()
    return out


@pytest.mark.parametrize("name,over", [
    ("sc_heis_afm_32", dict(sc="6 6 6")),
    ("sc_ising_16", dict(sc="6 6 6", para=".T. 1")),
])
def test_graph_replays_equal_eager_launches(name, over):
    say(f"graphs {name}")
    # period lcm(10, 1, 11) = 110: periods before and after mcs_eq = 100, and a tail
    chunks = [220, 115, 330, 7]
    a = _run(name, True, chunks, **over)
    b = _run(name, False, chunks, **over)
    assert a["sc"] == b["sc"]
    assert np.array_equal(a["spins"], b["spins"])
    assert np.array_equal(a["mom"], b["mom"])
    for x, y in zip(a["pt"], b["pt"]):
        assert np.array_equal(x, y)
    assert a["mom"][:, E.MOM_N].sum() > 0
