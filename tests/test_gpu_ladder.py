"""What the learned tempering ladder gains over the linear one.

Pyrochlore tempering (2^3 cubic cells x 16 sites, J1 = 1, J2 = 0.05 meV) on
20 -> 1 K (T/J 1.7 -> 0.086) with 16 temperatures and 16 independent ladders.
Below that range uniform Marsaglia proposals are accepted at a rate of about
kT/4J or less, the chains' energies hardly move between swaps, and no label
completes a round trip on any ladder within the run; there the gain shows only
in the swap acceptance (the benchmark reports both).

Both runs spend their first 2000 MCS in four training windows, the first on
the linear ladder, then three validation windows of 250 MCS (plain, learned,
plain).  The plain run stays on the linear ladder; the learned run retunes its
slots after each window to the ladder placed by the device MLP (trained on the
measured <e> and the acceptance-consistent C) and keeps it for production
where it beat the linear ladder in validation.  Both then run the same number
of production MCS (97250, about 8800 swap rounds: with 16 temperatures a label
needs a few hundred rounds per round trip even at uniform acceptance, since
its walk along the ladder is close to diffusive at acceptances of 0.1-0.4).
The linear ladder's large relative steps at the cold end
make the coldest pairs almost never swap (for 128 spins with C ~ N the coldest
pair's acceptance is erfc(dT sqrt(C) / 2T) ~ 1e-10); the learned ladder must
give a more uniform swap acceptance (lower coefficient of variation, higher
minimum) and more completed cold->hot->cold round trips.  Swaps pass the exact
acceptance test in both runs, so the sampled distributions are unchanged.

What the network adds: the same protocol with the ladder placed straight from
the length map (the thermodynamic length s(T) of the measured acceptances,
piecewise linear, no network) against the MLP placement (the network's
smooth fit of the same s(T) over ln T).  Over six seeds both were kept in
five and gave 165 +- 30 (map) and 162 +- 30 (MLP) round trips, against none
on the linear ladder: the network matches the map, it does not improve on
it at this size.  The test runs three seeds on which both are kept.
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
    assert lrn.source == "learned"
    assert lrn.round_trips > 0
    assert lrn.round_trips > lin.round_trips


def test_mlp_placement_against_the_length_map_alone():
    say("ladder ablation: MLP placement against the length map")
    from ether_amd.driver import make_engine
    from ether_amd.inputs import fortran_nint
    m = workloads.model("pyro_pt_16", sc="2 2 2", temp="20 1 1.26666666666667",
                        repeat="16", mcs="100000 3000")
    t = m.tags
    temps = t.temperatures()
    cv = lambda r: r.acc_rate.std() / r.acc_rate.mean()
    trips = {"linear": [], "mlp": [], "map": []}
    for seed in (1000, 1001, 1002):
        runs = {}
        for name, learn, pl in (("linear", False, 0), ("mlp", True, ladder.PLACEMENT_MLP),
                                ("map", True, ladder.PLACEMENT_MAP)):
            eng = make_engine(m, len(temps) * t.repeat, seed=seed)
            try:
                runs[name] = ladder.adaptive_run(eng, m, temps, t.repeat, t.tmcs, fortran_nint(t.coa),
                                                 t.exchange_interval, learn=learn, train_mcs=2000, placement=pl)
            finally:
                eng.close()
            r = runs[name]
            say(f"seed {seed} {name:6s}: {r.source} trips {r.round_trips} cost {r.cost:.1f} "
                f"acc min {r.acc_rate.min():.3f} cv {cv(r):.3f}")
            trips[name].append(r.round_trips)
        # both placements pass validation and beat the linear ladder
        for name in ("mlp", "map"):
            assert runs[name].source == "learned"
            assert runs[name].round_trips > runs["linear"].round_trips
            assert runs[name].acc_rate.min() > runs["linear"].acc_rate.min()
            assert cv(runs[name]) < cv(runs["linear"])
    # the network keeps what the map measured: its round trips within 15% of the map's
    assert np.mean(trips["mlp"]) >= 0.85 * np.mean(trips["map"])


def test_rejected_learned_ladder_repeats_the_linear_run():
    """Validation windows of one swap round cannot tell the ladders apart:
    the learned ladder is rejected, the context goes back to its state before
    training and runs the windows on the linear ladder, and the whole run
    (training, validation, production) is then the linear run bit for bit."""
    say("ladder: rejection repeats the linear run")
    m = workloads.model("pyro_pt_16", sc="2 2 2", temp="20 1 1.26666666666667",
                        repeat="4", mcs="1500 1000")
    lin, lrn = ladder.compare_ladders(m, train_mcs=800, validate_mcs=11)
    assert lin.source == "plain" and lrn.source == "plain"
    assert np.array_equal(lrn.temps, lin.temps)
    assert lrn.round_trips == lin.round_trips
    assert np.array_equal(lrn.acc_rate, lin.acc_rate)
    assert np.array_equal(lrn.metro_acc, lin.metro_acc)
