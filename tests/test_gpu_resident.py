"""Ising sweeps with the lattice resident in LDS against the colour phases
launched one by one (ether_set_resident(ctx, 0)): one workgroup per word of
four replicas (k_ising_resident, mode 2) and one per replica
(k_ising_resident1, mode 3).

Both resident kernels make every update with the same Philox block, the same
down counts and the same integer threshold as k_ising_sweep, so a whole
schedule (sweeps, tempering swaps every few MCS, samples, graphs) must leave
bit-identical spins, moments, swap counters and round trips; and
ether_sweep's per-sweep acceptance must be equal.  The oracle parity of the
Ising sweep (tests/test_gpu_parity.py, ether_sweep) runs on the resident path
too, since it is the default wherever a lattice fits.  Cases: one coupling
class (8^3, and the 16^3 x 190 workload), two classes with Zeeman (pyrochlore
two shells), three classes (simple cubic with three shells: the kernels'
path for any number of classes).
"""
import numpy as np
import pytest

from ether_amd import engine as E
from ether_amd import workloads
from ether_amd.driver import make_engine
from tests.helpers import pair, say, three_shell_ising


def _three_shell_ising(L, R):
    """Simple-cubic Ising with three coupling classes (26 neighbours per site)."""
    m = three_shell_ising(L)
    return m, make_engine(m, R, seed=1992)

pytestmark = pytest.mark.gpu


def _ising_start(m, engs, R, seed):
    rng = np.random.default_rng(seed)
    for r in range(R):
        S = np.zeros((m.n_sites, 3))
        S[:, 2] = np.where(rng.random(m.n_sites) < 0.5, 1.0, -1.0)
        for e in engs:
            e.set_spins(r, S)


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("name,over,R,ladder", [
    ("sc_ising_16", dict(sc="8 8 8"), 70, 10),
    ("sc_ising_16", dict(), 190, 19),                       # the 16^3 workload
    ("pyro_pt_16", dict(sc="2 2 2", model="ising", para=".T. 1", pt=".F. 11",
                        zeeman=".T. 0 0 0.3"), 130, 13),   # two shells (two classes), Zeeman, 4 colours
    ("sc_3shell", dict(), 60, 12),                          # three classes, 26 neighbours, 8 colours
])
def test_resident_schedule_equals_phase_kernels(name, over, R, ladder, mode):
    say(f"resident (mode {mode}) vs phase kernels {name} {over} R={R}")
    if name == "sc_3shell":
        m, res = _three_shell_ising(6, R)
    else:
        m, res, _ = pair(name, R, **over)
    res.set_resident(mode)
    phase = make_engine(m, R, seed=1992)
    phase.set_resident(False)
    temps = np.tile(np.linspace(2.0, 6.0, ladder) if m.tags.para else np.linspace(10.0, 100.0, ladder),
                    R // ladder + 1)[:R]
    for e in (res, phase):
        e.set_betas(m.betas(temps))
        e.set_ladders(list(range(0, R - R % ladder + 1, ladder)))
    _ising_start(m, (res, phase), R, 3)
    sc = [0, 0]
    # equilibration then production: samples every 5 MCS, swaps every 3 MCS,
    # graphs for whole periods, launches one by one for the rest
    for k, e in enumerate((res, phase)):
        sc[k] = e.run(1, 37, 20, 5, 0, 0.0, 3, 0)
        sc[k] = e.run(38, 23, 20, 5, 0, 0.0, 3, sc[k])
    assert sc[0] == sc[1]
    for r in range(R):
        a, _ = res.get_spins(r)
        b, _ = phase.get_spins(r)
        assert np.array_equal(a, b), f"replica {r}"
    assert np.array_equal(res.moments(), phase.moments())
    for x, y in zip(res.pt_state(), phase.pt_state()):
        assert np.array_equal(x, y)
    pa, ba = res.energy_parts()
    pb, bb = phase.energy_parts()
    assert np.array_equal(pa, pb) and np.array_equal(ba, bb)
    # ether_sweep: the last sweep's acceptance per replica
    aa = res.sweep(4, want_accept=True)
    ab = phase.sweep(4, want_accept=True)
    assert np.array_equal(aa, ab)
    for r in (0, R // 2, R - 1):
        assert np.array_equal(res.get_spins(r)[0], phase.get_spins(r)[0])
    res.close()
    phase.close()


def test_resident_mode_is_checked():
    R = 8
    m, eng, _ = pair("sc_ising_16", R, sc="4 4 4")
    with pytest.raises(E.EtherError):
        eng.set_resident(4)
    eng.close()


def test_resident_launch_is_timed_as_its_own_kind():
    say("resident timing kind")
    R = 64
    m, eng, _ = pair("sc_ising_16", R, sc="8 8 8")
    eng.set_betas(m.betas(np.linspace(10.0, 100.0, R)))
    eng.set_timing(True)
    eng.run(1, 20, 0, 10)
    ms, n = eng.timing(E.TIME_RESIDENT)
    ms0, n0 = eng.timing(E.TIME_SWEEP)
    assert n == 20 and ms > 0.0 and n0 == 0
    eng.set_timing(False)
    eng.close()
