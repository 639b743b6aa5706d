"""What the learned tempering ladder gains over the linear one.

A small pyrochlore tempering run (2^3 cubic cells x 16 sites) on the
workload's 1 -> 0.05 K range.  The linear ladder crowds the cold end with
large relative steps (dT/T), so swaps there are rarely accepted; the learned
ladder (device MLP trained on measured <e>, C_v, then equal sqrt(C_v)/T
length) must give a more uniform swap acceptance and more completed
cold->hot->cold round trips in the same number of MCS.  The exact swap test
is unchanged, so per-temperature energies from both ladders agree where the
ladders share a temperature (the two end points).
"""
import numpy as np
import pytest

from ether_amd import ladder, workloads
from ether_amd.driver import make_engine
from ether_amd.inputs import fortran_nint
from tests.helpers import say

pytestmark = pytest.mark.gpu


def test_learned_ladder_beats_linear_on_round_trips():
    say("ladder gain")
    m = workloads.model("pyro_pt_16", sc="2 2 2", temp="1 0.05 0.0633333333333333",
                        repeat="16")
    t = m.tags
    temps = t.temperatures()
    assert len(temps) == 16
    rep = t.repeat
    eng = make_engine(m, len(temps) * rep)
    coa = fortran_nint(t.coa)
    lin = ladder.run_ladder(eng, m, temps, rep, 6000, coa, t.exchange_interval)
    e, cv = ladder.measure_on_ladder(eng, m, temps, rep, 2000, coa, t.exchange_interval)
    T_new, loss = ladder.learn_ladder(eng, temps, e, cv)
    lrn = ladder.run_ladder(eng, m, T_new, rep, 6000, coa, t.exchange_interval)
    eng.close()
    say(f"linear: trips {lin.round_trips} acc {lin.acc_rate.min():.3f}..{lin.acc_rate.max():.3f}")
    say(f"learned: trips {lrn.round_trips} acc {lrn.acc_rate.min():.3f}..{lrn.acc_rate.max():.3f}")
    say(f"ladder {np.round(T_new, 4)}")
    assert T_new[0] == pytest.approx(1.0) and T_new[-1] == pytest.approx(0.05)
    assert np.all(np.diff(T_new) < 0)
    assert lrn.acc_rate.min() > lin.acc_rate.min()
    assert lrn.acc_rate.std() < lin.acc_rate.std()
    assert lrn.round_trips > lin.round_trips
