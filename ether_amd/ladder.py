"""Learned parallel-tempering ladder.

The reference's ladder is linear in T (src/ether.f90:79-84); an adaptive
ladder is only a commented stub (src/startup.f90:115-122).  Here a small MLP
surrogate of <e>(T) and log C(T), trained by the library's own HIP kernels on
measurements from the first MCS of the run, places the temperatures so that
the thermodynamic length  int sqrt(C)/T dT  is equal between neighbours.  For
Gaussian energy distributions the swap acceptance of a neighbour pair is
erfc(dT sqrt(C)/(2T)), so equal length means uniform acceptance.  Uniform
acceptance minimises the round-trip time of the deterministic even/odd swap
schedule that the reference uses (swap_count parity,
src/parallel_tempering.f90:50-62).

Which C.  The fluctuation estimator beta^2 (<e^2> - <e>^2) N over a short
window is biased by relaxation drift wherever the chains are still far from
equilibrium (at T/J << 1 uniform Marsaglia proposals are almost never
accepted).  The C that actually sets the swap acceptance is the one that
reproduces the measured acceptance a_k of each neighbour pair through the same
Gaussian formula:  sqrt(C(T_mid)) / T_mid = 2 erfcinv(a_k) / dT_k.  These
densities are kept on a grid over the temperature range (the length map); the
surrogate is trained on log C from that map and on <e>.

Protocol (ether_ladder_train in the library, so that the Fortran driver runs
the same one): the first train_mcs MCS are split into windows.  The first runs
on the linear ladder; after each window the map takes the intervals the window
measured, the MLP is retrained and the temperatures of the slots are retuned in
place (every replica keeps its state and slot), so a pair whose acceptance was
too small to measure on one ladder (a lower bound only) is measured on the
next.  Three validation windows follow (plain, learned, plain): the learned
ladder is kept only where it beats the plain one on both the round-trip cost
and the lowest pair acceptance.  Every swap passes the exact acceptance test
throughout, so the sampled distribution at each temperature is unchanged; only
the production window after the validation is timed and its round trips
counted.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import engine as E
from .driver import make_engine
from .inputs import Model, fortran_nint


@dataclass
class LadderRun:
    temps: np.ndarray          # (n,) production ladder, hottest first
    wall_s: float              # wall time of the production window
    mcs: int                   # MCS in the production window
    round_trips: int           # completed cold->hot->cold trips, all replicas
    acc_rate: np.ndarray       # (n-1,) swap acceptance between slot k and k+1
    updates: float
    mlp_loss: Optional[float] = None
    metro_acc: Optional[np.ndarray] = None   # (n,) Metropolis acceptance per slot, production
    source: str = "plain"                     # which ladder ran the production window
    train_cost: Optional[float] = None        # its round-trip cost measured in validation

    @property
    def cost(self) -> float:
        """Round-trip cost Lambda of the production window."""
        return round_trip_cost(self.acc_rate, np.full(len(self.acc_rate), 1e9))

    @property
    def round_trips_per_s(self) -> float:
        return self.round_trips / self.wall_s if self.wall_s > 0 else 0.0

    @property
    def updates_per_s(self) -> float:
        return self.updates / self.wall_s if self.wall_s > 0 else 0.0


def _pair_acceptance(eng: E.Engine, n: int, repeat: int):
    _, acc, att, trips = eng.pt_state()
    a = np.zeros(n - 1)
    tot = np.zeros(n - 1)
    for k in range(n - 1):
        num = sum(int(acc[g * n + k]) for g in range(repeat))
        den = sum(int(att[g * n + k]) for g in range(repeat))
        a[k] = num / den if den else 0.0
        tot[k] = den
    return a, tot, int(trips.sum())


def round_trip_cost(acc, attempts) -> float:
    """Lambda = sum_k r_k / (1 - r_k) over the neighbour pairs of a ladder,
    r_k the measured swap rejection rate (a pair with no accepted swap counts
    as half an acceptance).  Under the even/odd swap schedule a label needs
    about 2 + 2 Lambda swap rounds per round trip (Syed, Bouchard-Cote,
    Deligiannidis and Doucet, J. R. Stat. Soc. B 84 (2022), non-reversible
    parallel tempering), so the smaller Lambda the more round trips."""
    att = np.maximum(np.asarray(attempts, np.float64), 1.0)
    a = np.clip(np.asarray(acc, np.float64), 0.5 / att, 1.0)
    return float(np.sum((1.0 - a) / a))


PLACEMENT_MLP, PLACEMENT_MAP = 0, 1


@dataclass
class Training:
    ladder: np.ndarray         # the ladder set for production (hottest first)
    next_step: int             # first MCS after the training and validation windows
    swap_count: int
    source: str                # "plain" or "learned"
    cost: float                # its round-trip cost Lambda in validation
    cost_plain: float = float("nan")
    min_acc: float = float("nan")
    min_acc_plain: float = float("nan")
    loss: Optional[float] = None


def train_ladder(eng: E.Engine, model: Model, temps, repeat: int, coa: int, pt_interval: int,
                 train_mcs: int, *, windows: int = 4, learn: bool = True, first: int = 1,
                 swap_count: int = 0, ovrr_iv: int = 0, ovrr_p: float = 0.0,
                 validate_mcs: Optional[int] = None, placement: int = PLACEMENT_MLP) -> Training:
    """The training windows of a tempering run (ether_train_ladder in the
    library, which the Fortran driver calls too): train_mcs MCS from step
    `first` in `windows` windows, the first on `temps` (the plain ladder); with
    learn=True the ladder is re-placed after each window from the measured
    swap acceptances (by the device MLP, or straight from the length map with
    placement=PLACEMENT_MAP), and the slots are retuned in place.  Then three
    validation windows of validate_mcs MCS (default train_mcs / 8): plain,
    learned, plain; the learned ladder is kept only if its round-trip cost is
    lower by more than twice the standard error of the difference and its
    lowest pair acceptance higher than the plain ladder's; otherwise the
    context returns to its state at the start and runs the same windows on
    the plain ladder, so that production continues as with learn=False.  learn=False runs the same
    windows on `temps`.  Every replica keeps its state and slot; the windows
    follow the run's schedule (OVRR every ovrr_iv MCS with ovrr_p passes)."""
    temps = np.asarray(temps, np.float64)
    val = train_mcs // 8 if validate_mcs is None else validate_mcs
    T, sc, rep = eng.ladder_train(temps, model.kb, first, train_mcs, windows, val, coa, ovrr_iv,
                                  ovrr_p, pt_interval, learn, placement, swap_count=swap_count)
    kept = bool(rep[4])
    return Training(T, first + int(rep[6]), sc, "learned" if kept else "plain",
                    float(rep[1] if kept else rep[0]), cost_plain=float(rep[0]), min_acc=float(rep[3]),
                    min_acc_plain=float(rep[2]), loss=None if np.isnan(rep[5]) else float(rep[5]))


def adaptive_run(eng: E.Engine, model: Model, temps, repeat: int, mcs: int, coa: int,
                 pt_interval: int, *, learn: bool, train_mcs: int = 2000, windows: int = 4,
                 validate_mcs: Optional[int] = None, placement: int = PLACEMENT_MLP) -> LadderRun:
    """Tempering run of mcs MCS from random states: train_ladder (training and
    validation windows) first, then the remaining MCS are the timed
    production window on the ladder it set."""
    temps = np.asarray(temps, np.float64)
    n = len(temps)
    eq = model.tags.tmcs_eq
    eng.set_betas(np.tile(model.betas(temps), repeat))
    eng.init_spins(E.INIT_RANDOM)
    tr = train_ladder(eng, model, temps, repeat, coa, pt_interval, train_mcs, windows=windows,
                      learn=learn, validate_mcs=validate_mcs, placement=placement)
    cur, first, sc = tr.ladder, tr.next_step, tr.swap_count
    # production window
    eng.set_ladders([k * n for k in range(repeat + 1)])
    eng.reset_moments()
    eng.synchronize()
    rest = mcs - (first - 1)
    t0 = time.perf_counter()
    sc = eng.run(first, rest, eq, coa, 0, 0.0, pt_interval, sc)
    eng.synchronize()
    wall = time.perf_counter() - t0
    acc, _, trips = _pair_acceptance(eng, n, repeat)
    mom = eng.moments()
    metro = np.array([np.mean([mom[g * n + k, E.MOM_ACC] for g in range(repeat)]) / rest
                      for k in range(n)])
    return LadderRun(cur, wall, rest, trips, acc, float(eng.R) * model.n_sites * rest,
                     mlp_loss=tr.loss, metro_acc=metro, source=tr.source, train_cost=tr.cost)


def compare_ladders(model: Model, *, train_mcs: int = 2000, mcs: Optional[int] = None,
                    repeat: Optional[int] = None, placement: int = PLACEMENT_MLP,
                    validate_mcs: Optional[int] = None):
    """The linear ladder against the learned ladder on the same model and seed,
    each on a context of its own built alike (the same random start): where
    validation rejects the learned ladder, its run repeats the linear one
    exactly.  Returns (linear LadderRun, learned LadderRun)."""
    t = model.tags
    temps = t.temperatures()
    rep = t.repeat if repeat is None else repeat
    mcs = t.tmcs if mcs is None else mcs
    coa = fortran_nint(t.coa)

    def run(learn):
        eng = make_engine(model, len(temps) * rep)
        try:
            return adaptive_run(eng, model, temps, rep, mcs, coa, t.exchange_interval, learn=learn,
                                train_mcs=train_mcs, placement=placement, validate_mcs=validate_mcs)
        finally:
            eng.close()

    return run(False), run(True)


def plain_run(model: Model, *, train_mcs: int = 2000, mcs: Optional[int] = None,
              repeat: Optional[int] = None) -> LadderRun:
    """The plain GPU form: one context per ladder (repeat), each with the
    linear ladder and launched on its own (one launch per lattice per kernel,
    its own graphs), the contexts run one after another by the host.  The
    same MCS before production as adaptive_run (train_mcs plus three
    validation windows of train_mcs / 8); the production windows of all
    ladders are timed together."""
    t = model.tags
    temps = np.asarray(t.temperatures(), np.float64)
    n = len(temps)
    rep = t.repeat if repeat is None else repeat
    mcs = t.tmcs if mcs is None else mcs
    coa = fortran_nint(t.coa)
    eq = t.tmcs_eq
    pre = train_mcs + 3 * (train_mcs // 8)
    engines = [make_engine(model, n, seed=t.seed + 7919 * k) for k in range(rep)]
    try:
        for eng in engines:
            eng.set_betas(model.betas(temps))
            eng.set_ladders([0, n])
            eng.init_spins(E.INIT_RANDOM)
            eng.run(1, pre, eq, coa, 0, 0.0, t.exchange_interval, 0)
        scs = []
        for eng in engines:
            eng.set_ladders([0, n])
            eng.reset_moments()
            eng.synchronize()
        t0 = time.perf_counter()
        rest = mcs - pre
        for eng in engines:
            scs.append(eng.run(pre + 1, rest, eq, coa, 0, 0.0, t.exchange_interval, 0))
        for eng in engines:
            eng.synchronize()
        wall = time.perf_counter() - t0
        acc = np.zeros(n - 1)
        att = np.zeros(n - 1)
        trips = 0
        metro = np.zeros(n)
        for eng in engines:
            _, a, b, tr = eng.pt_state()
            acc += a[: n - 1]
            att += b[: n - 1]
            trips += int(tr.sum())
            metro += eng.moments()[:, E.MOM_ACC] / rest / rep
        return LadderRun(temps, wall, rest, trips, acc / np.maximum(att, 1),
                         float(n * rep) * model.n_sites * rest, metro_acc=metro, source="plain")
    finally:
        for eng in engines:
            eng.close()
