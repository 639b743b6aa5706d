"""What the learned tempering ladder gains over the linear one.

Pyrochlore tempering (2^3 cubic cells x 16 sites, J1 = 1, J2 = 0.05 meV) on
20 -> 1 K (T/J 1.7 -> 0.086) with 16 temperatures and 16 independent ladders.
Below that range uniform Marsaglia proposals are accepted at a rate of about
kT/4J or less, the chains' energies hardly move between swaps, and no label
completes a round trip on any ladder within the run; there the gain shows only
in the swap acceptance (the benchmark reports both).

Both runs spend their first 2000 MCS in four training windows, the first on
the linear ladder.  The plain run stays on it; the learned run retunes its
slots after each window to the ladder placed by the device MLP (trained on the
measured <e> and the acceptance-consistent C).  Both then run the same number
of production MCS (98000, 8900 swap rounds: with 16 temperatures a label
needs a few hundred rounds per round trip even at uniform acceptance, since
its walk along the ladder is close to diffusive at acceptances of 0.1-0.4).
The linear ladder's large relative steps at the cold end
make the coldest pairs almost never swap (for 128 spins with C ~ N the coldest
pair's acceptance is erfc(dT sqrt(C) / 2T) ~ 1e-10); the learned ladder must
give a more uniform swap acceptance (lower coefficient of variation, higher
minimum) and more completed cold->hot->cold round trips.  Swaps pass the exact
acceptance test in both runs, so the sampled distributions are unchanged.
"""
import numpy as np
import pytest

from ether_amd import ladder, workloads
from tests.helpers import say

pytestmark = pytest.mark.gpu


def test_learned_ladder_beats_linear_on_round_trips():
    say("ladder gain")
    m = workloads.model("pyro_pt_16", sc="2 2 2", temp="20 1 1.26666666666667",
                        repeat="16", mcs="100000 3000")
    assert len(m.tags.temperatures()) == 16
    lin, lrn = ladder.compare_ladders(m, train_mcs=2000)
    cv = lambda r: r.acc_rate.std() / r.acc_rate.mean()
    say(f"linear : trips {lin.round_trips} acc {lin.acc_rate.min():.3f}..{lin.acc_rate.max():.3f} cv {cv(lin):.3f}")
    say(f"learned: trips {lrn.round_trips} acc {lrn.acc_rate.min():.3f}..{lrn.acc_rate.max():.3f} cv {cv(lrn):.3f}")
    say(f"ladder {np.round(lrn.temps, 4)}")
    T = lrn.temps
    assert T[0] == pytest.approx(20.0) and T[-1] == pytest.approx(1.0)
    assert np.all(np.diff(T) < 0)
    assert cv(lrn) < cv(lin)
    assert lrn.acc_rate.min() > lin.acc_rate.min()
    assert lrn.round_trips > 0
    assert lrn.round_trips > lin.round_trips
