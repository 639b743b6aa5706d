"""Launch shape and autotuner: results must not depend on them.

Every update draws from its own (site, replica, sweep) Philox counter and
reads only spins of other colour classes, so the sites-per-wave and XCD
mapping of a colour phase change which wavefront does an update, never its
arithmetic: spins after K sweeps are bit-identical for every shape.  The
autotuner must leave the spins, the acceptance counters and the sweep counter
as it found them, so that the run that follows is the run without it.
"""
import numpy as np
import pytest

from ether_amd import engine as E
from ether_amd import workloads
from ether_amd.driver import attach_neighbours, make_engine
from tests.helpers import pair, random_xyz, say

pytestmark = pytest.mark.gpu


def _start(m, eng, R, seed):
    rng = np.random.default_rng(seed)
    for r in range(R):
        if m.tags.ising:
            S = np.zeros((m.n_sites, 3))
            S[:, 2] = np.where(rng.random(m.n_sites) < 0.5, 1.0, -1.0)
            eng.set_spins(r, S)
        else:
            S = random_xyz(rng, m.n_sites)
            eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))


def _spins(eng, R):
    return np.stack([eng.get_spins(r)[0] for r in range(R)])


@pytest.mark.parametrize("name,over,R,lanes", [
    ("sc_heis_afm_32", dict(sc="8 8 8"), 70, [0]),
    ("sc_ising_16", dict(sc="8 8 8"), 300, [0]),
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 24, [64, 32, 0, 25]),      # 25 lanes: 2 sites per step
    ("tri_afm_96", dict(sc="12 12 1"), 13, [64, 16, 32, 0, 21, 13]),         # ANGLE, 3 colours; 13 and 21 lanes
    ("sc_heis_afm_32", dict(sc="6 6 6"), 20, [64, 32, 0, 20, 8 * 3]),       # 20 lanes: 3 sites per step
    ("sc_heis_afm_32", dict(sc="4 4 4"), 8, [0, 64, 8, 11]),                # 8 lanes: 8 sites per step
    ("sc_ising_16", dict(sc="6 6 6"), 30, [64, 8, 16, 32]),
])
def test_sweeps_do_not_depend_on_launch_shape(name, over, R, lanes):
    """Sites per wave, the XCD map and lanes per site (several sites per wave
    for few replicas) change which lanes do an update, never its result."""
    say(f"launch shapes {name} R={R} lanes {lanes}")
    ref = None
    shapes = [(4, 1, lanes[0]), (1, 0, lanes[-1]), (2, 1, lanes[0]), (8, 0, lanes[-1]), (16, 1, lanes[0]),
              (4, 2, lanes[0]), (1, 2, lanes[-1]), (16, 2, lanes[0])]
    shapes += [(4, 1, ln) for ln in lanes]
    for spw, xcd, ln in shapes:
        m, eng, _ = pair(name, R, **over)
        eng.set_lanes(ln)
        eng.set_betas(m.betas(np.linspace(5.0, 50.0, R)))
        _start(m, eng, R, 3)
        eng.set_launch(spw, xcd)
        acc = eng.sweep(3, want_accept=True)
        got = (_spins(eng, R), acc)
        eng.close()
        if ref is None:
            ref = got
        else:
            assert np.array_equal(got[0], ref[0]), \
                f"spins differ with sites_per_wave={spw} xcd_map={xcd} lanes={ln}"
            assert np.array_equal(got[1], ref[1])


@pytest.mark.parametrize("name,over,R", [
    ("sc_heis_afm_32", dict(sc="8 8 8"), 70),
    ("sc_ising_16", dict(sc="8 8 8"), 300),
])
def test_autotune_restores_state_and_changes_no_result(name, over, R):
    say(f"autotune {name} R={R}")
    runs = []
    for tune in (False, True):
        m, eng, _ = pair(name, R, **over)
        eng.set_betas(m.betas(np.linspace(5.0, 50.0, R)))
        _start(m, eng, R, 4)
        eng.sweep(2)
        before = _spins(eng, R)
        mom_before = eng.moments()
        if tune:
            spw, xcd, table = eng.autotune(reps=3)
            say(f"  chose sites_per_wave={spw} xcd_map={xcd}; ms per sweep {np.round(table, 4)}")
            assert spw in E.Engine.AUTOTUNE_SPW and xcd in (0, 1, 2)
            assert table.shape == (3, len(E.Engine.AUTOTUNE_SPW))
            assert np.all(table > 0)
            # the fastest, or the most sites per wave among those within 1% of it
            chosen = table[xcd][E.Engine.AUTOTUNE_SPW.index(spw)]
            assert chosen <= 1.01 * table.min()
            near = [E.Engine.AUTOTUNE_SPW[i] for x in range(3) for i in range(table.shape[1])
                    if table[x][i] <= 1.01 * table.min()]
            assert spw == max(near)
            assert np.array_equal(_spins(eng, R), before)
            assert np.array_equal(eng.moments(), mom_before)
        eng.sweep(3)
        runs.append(_spins(eng, R))
        eng.close()
    assert np.array_equal(runs[0], runs[1])


def test_event_timing_counts_every_colour_phase():
    """One event pair brackets the colour phases of a sweep; the reported
    launch count is still one per phase."""
    m = workloads.model("sc_heis_afm_32", sc="6 6 6")
    attach_neighbours(m)
    eng = make_engine(m, 40, seed=3)
    eng.set_betas(m.betas(np.linspace(30, 3, 40)))
    eng.init_spins(E.INIT_RANDOM)
    nc = eng.layout()[0]
    eng.set_timing(True)
    eng.run(1, 12, 1000, 10, 1, 1.0, 0, 0)
    ms, n = eng.timing(E.TIME_SWEEP)
    eng.set_timing(False)
    assert n == 12 * nc and ms > 0.0
    # with samples every 4 MCS, the sweeps before them book the measurement
    # (their own timing kind) and only the finalize is timed as measurement
    eng.set_timing(True)
    eng.run(13, 12, 0, 4, 1, 1.0, 0, 0)
    assert eng.timing(E.TIME_SWEEP)[1] == 9 * nc
    assert eng.timing(E.TIME_SWEEP_MEAS)[1] == 3 * nc
    assert eng.timing(E.TIME_MEASURE)[1] == 3
    eng.set_timing(False)
    eng.close()
