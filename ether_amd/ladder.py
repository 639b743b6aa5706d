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
densities are kept on a grid over the temperature range (LengthMap); the
surrogate is trained on log C from that map and on <e>.

Protocol (adaptive_run): the first train_mcs MCS of the equilibration run are
split into windows.  The first runs on the linear ladder; after each window
the map takes the intervals the window measured, the MLP is retrained and the
temperatures of the slots are retuned in place (every replica keeps its state
and slot), so a pair whose acceptance was too small to measure on one ladder
(a lower bound only) is measured on the next.  The remaining MCS run on the
last ladder.  Every swap passes the exact acceptance test throughout, so the
sampled distribution at each temperature is unchanged; only the production
window after the last retune is timed and its round trips counted.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
from scipy.special import erfcinv

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
    train_e: Optional[np.ndarray] = None
    train_c: Optional[np.ndarray] = None
    train_acc: Optional[np.ndarray] = None
    mlp_loss: Optional[float] = None
    metro_acc: Optional[np.ndarray] = None   # (n,) Metropolis acceptance per slot, production
    source: str = "plain"                     # which ladder ran the production window
    train_cost: Optional[float] = None        # its round-trip cost measured in training

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


def length_density(temps, acc, attempts):
    """Thermodynamic-length density of each interval of a ladder, from its
    swap acceptance.  For Gaussian energy distributions a = erfc(d/2) with
    d = int sqrt(C)/T dT over the interval, so lambda = 2 erfcinv(a) / |dT|
    (sqrt(C)/T at the midpoint for smooth C).  A pair that accepted no swap
    only bounds d from below; it is flagged bound=True."""
    temps = np.asarray(temps, np.float64)
    att = np.maximum(np.asarray(attempts, np.float64), 1.0)
    acc = np.asarray(acc, np.float64)
    bound = acc * att < 0.5
    a = np.clip(acc, 0.5 / att, 0.999)
    lam = 2.0 * erfcinv(a) / np.abs(np.diff(temps))
    return lam, bound


class LengthMap:
    """sqrt(C)/T on a fixed grid over the temperature range, refined window by
    window.  A measured interval replaces what earlier windows said on its
    span; an interval with no accepted swap only raises the value there to
    its lower bound."""

    def __init__(self, t_lo: float, t_hi: float, n: int = 64):
        self.T = np.linspace(t_lo, t_hi, n)
        self.lam = np.full(n, np.nan)

    def add(self, temps, lam, bound) -> None:
        temps = np.asarray(temps, np.float64)
        lo = np.minimum(temps[:-1], temps[1:])
        hi = np.maximum(temps[:-1], temps[1:])
        for k in range(len(lam)):
            sel = (self.T >= lo[k]) & (self.T <= hi[k])
            if bound[k]:
                self.lam[sel] = np.fmax(self.lam[sel], lam[k])
            else:
                self.lam[sel] = lam[k]

    def heat_capacity(self) -> np.ndarray:
        return (self.lam * self.T) ** 2


def learn_ladder(eng: E.Engine, temps, e, c, n_out: Optional[int] = None, iters: int = 4000,
                 lr: float = 3e-3, seed: int = 1992, t_range=None):
    """Train the device MLP on samples (T, e, log C) and place a new ladder of
    n_out temperatures on t_range (default: the span of temps)."""
    temps = np.asarray(temps, np.float64)
    n_out = n_out or len(temps)
    t_lo, t_hi = t_range if t_range is not None else (float(temps.min()), float(temps.max()))
    mid, half = 0.5 * (t_lo + t_hi), 0.5 * (t_hi - t_lo)
    x = ((temps - mid) / half).astype(np.float32)
    lc = np.log(np.maximum(c, 1e-12))
    y = np.stack([(e - e.mean()) / (e.std() + 1e-30), (lc - lc.mean()) / (lc.std() + 1e-30)],
                 axis=1).astype(np.float32)
    w = np.ones_like(y)
    eng.mlp_init(seed)
    loss = eng.mlp_train(x, y, w, iters, lr)
    T = eng.ladder_place(mid, half, float(lc.std() + 1e-30), t_lo, t_hi, n_out)
    return T, loss


def _window(eng: E.Engine, model: Model, temps, repeat: int, first: int, n_mcs: int, coa: int,
            pt_interval: int, sc: int, ovrr_iv: int = 0, ovrr_p: float = 0.0):
    """Run n_mcs MCS from step `first` with fresh swap counters and energy
    moments; returns (pair acceptance, attempts, <e> per slot, swap counter)."""
    n = len(temps)
    eng.set_ladders([k * n for k in range(repeat + 1)])
    eng.reset_moments()
    # every sampling step of the window counts (eq = first - 1) so that <e> is measured
    sc = eng.run(first, n_mcs, first - 1, coa, ovrr_iv, ovrr_p, pt_interval, sc)
    mom = eng.moments()
    a, att, _ = _pair_acceptance(eng, n, repeat)
    e = np.zeros(n)
    for g in range(repeat):
        for k in range(n):
            m = mom[g * n + k]
            e[k] += m[E.MOM_E] / max(m[E.MOM_N], 1.0) / repeat
    return a, att, e, sc


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


@dataclass
class Training:
    ladder: np.ndarray         # the ladder chosen for production (hottest first)
    next_step: int             # first MCS after the training windows
    swap_count: int
    source: str                # "plain" or "learned (window k)"
    cost: float                # its measured round-trip cost Lambda
    e: Optional[np.ndarray] = None
    c: Optional[np.ndarray] = None
    acc: Optional[np.ndarray] = None
    loss: Optional[float] = None


def train_ladder(eng: E.Engine, model: Model, temps, repeat: int, coa: int, pt_interval: int,
                 train_mcs: int, *, windows: int = 4, learn: bool = True, first: int = 1,
                 swap_count: int = 0, ovrr_iv: int = 0, ovrr_p: float = 0.0) -> Training:
    """The training windows of a tempering run: train_mcs MCS from step
    `first`, split into `windows` equal windows; each one measures the pair
    acceptance and <e> of the ladder it ran on.  The first window runs on
    `temps` (the plain ladder).  learn=True: after every window but the last
    the length map takes the window's measured intervals, the MLP is retrained
    on it (log C = 2 log(lambda T) on the map's grid, <e> of the window
    interpolated) and the slots are retuned to the ladder it places, so each
    window measures the intervals that the previous ladder was too coarse to
    measure.  The ladder with the smallest measured round-trip cost Lambda is
    then set for production; the plain ladder is the fallback when no learned
    ladder beats it (far from equilibrium the acceptance is not the Gaussian
    one the surrogate assumes).  learn=False: every window runs on `temps`.
    Every replica keeps its state and slot throughout; the windows follow the
    run's schedule (OVRR every ovrr_iv MCS with ovrr_p passes)."""
    temps = np.asarray(temps, np.float64)
    n = len(temps)
    t_range = (float(temps.min()), float(temps.max()))
    win = max(train_mcs // windows, 1)
    cur, sc = temps, swap_count
    lmap = LengthMap(*t_range)
    best, best_cost, source = temps, np.inf, "plain"
    a = e = c = None
    loss = None
    for k in range(windows):
        length = win if k < windows - 1 else train_mcs - win * (windows - 1)
        if length <= 0:
            break
        a, att, e, sc = _window(eng, model, cur, repeat, first, length, coa, pt_interval, sc,
                                ovrr_iv, ovrr_p)
        first += length
        cost = round_trip_cost(a, att)
        if cost < best_cost:
            best, best_cost, source = cur, cost, ("plain" if k == 0 else f"learned (window {k})")
        lam, bound = length_density(cur, a, att)
        lmap.add(cur, lam, bound)
        c = lmap.heat_capacity()
        if learn and k < windows - 1:
            o = np.argsort(cur)
            e_grid = np.interp(lmap.T, cur[o], e[o])
            cur, loss = learn_ladder(eng, lmap.T, e_grid, c, n_out=n, t_range=t_range)
            eng.retune_betas(np.tile(model.betas(cur), repeat))
    if learn and best is not cur:
        cur = best
        eng.retune_betas(np.tile(model.betas(cur), repeat))
    return Training(np.asarray(cur), first, sc, source if learn else "plain", best_cost,
                    e=e, c=c, acc=a, loss=loss)


def adaptive_run(eng: E.Engine, model: Model, temps, repeat: int, mcs: int, coa: int,
                 pt_interval: int, *, learn: bool, train_mcs: int = 2000,
                 windows: int = 4) -> LadderRun:
    """Tempering run of mcs MCS from random states: train_ladder over the
    first train_mcs MCS, then the remaining mcs - train_mcs MCS are the timed
    production window on the chosen ladder."""
    temps = np.asarray(temps, np.float64)
    n = len(temps)
    eq = model.tags.tmcs_eq
    eng.set_betas(np.tile(model.betas(temps), repeat))
    eng.init_spins(E.INIT_RANDOM)
    tr = train_ladder(eng, model, temps, repeat, coa, pt_interval, train_mcs, windows=windows,
                      learn=learn)
    cur, first, sc = tr.ladder, tr.next_step, tr.swap_count
    # production window
    eng.set_ladders([k * n for k in range(repeat + 1)])
    eng.reset_moments()
    eng.synchronize()
    t0 = time.perf_counter()
    rest = mcs - train_mcs
    sc = eng.run(first, rest, eq, coa, 0, 0.0, pt_interval, sc)
    eng.synchronize()
    wall = time.perf_counter() - t0
    acc, _, trips = _pair_acceptance(eng, n, repeat)
    mom = eng.moments()
    metro = np.array([np.mean([mom[g * n + k, E.MOM_ACC] for g in range(repeat)]) / rest
                      for k in range(n)])
    return LadderRun(cur, wall, rest, trips, acc, float(eng.R) * model.n_sites * rest,
                     train_e=tr.e, train_c=tr.c, train_acc=tr.acc, mlp_loss=tr.loss, metro_acc=metro,
                     source=tr.source, train_cost=tr.cost)


def compare_ladders(model: Model, *, train_mcs: int = 2000, mcs: Optional[int] = None,
                    repeat: Optional[int] = None):
    """The linear ladder against the learned ladder on the same model and seed.
    Returns (linear LadderRun, learned LadderRun)."""
    t = model.tags
    temps = t.temperatures()
    rep = t.repeat if repeat is None else repeat
    mcs = t.tmcs if mcs is None else mcs
    coa = fortran_nint(t.coa)
    eng = make_engine(model, len(temps) * rep)
    try:
        lin = adaptive_run(eng, model, temps, rep, mcs, coa, t.exchange_interval, learn=False,
                           train_mcs=train_mcs)
        lrn = adaptive_run(eng, model, temps, rep, mcs, coa, t.exchange_interval, learn=True,
                           train_mcs=train_mcs)
    finally:
        eng.close()
    return lin, lrn


def plain_run(model: Model, *, train_mcs: int = 2000, mcs: Optional[int] = None,
              repeat: Optional[int] = None) -> LadderRun:
    """The plain GPU form: one context per ladder (repeat), each with the
    linear ladder and launched on its own (one launch per lattice per kernel,
    its own graphs), the contexts run one after another by the host.  Same
    schedule and windows as adaptive_run(learn=False); the production windows
    of all ladders are timed together."""
    t = model.tags
    temps = np.asarray(t.temperatures(), np.float64)
    n = len(temps)
    rep = t.repeat if repeat is None else repeat
    mcs = t.tmcs if mcs is None else mcs
    coa = fortran_nint(t.coa)
    eq = t.tmcs_eq
    engines = [make_engine(model, n, seed=t.seed + 7919 * k) for k in range(rep)]
    try:
        for eng in engines:
            eng.set_betas(model.betas(temps))
            eng.set_ladders([0, n])
            eng.init_spins(E.INIT_RANDOM)
            eng.run(1, train_mcs, eq, coa, 0, 0.0, t.exchange_interval, 0)
        scs = []
        for eng in engines:
            eng.set_ladders([0, n])
            eng.reset_moments()
            eng.synchronize()
        t0 = time.perf_counter()
        rest = mcs - train_mcs
        for eng in engines:
            scs.append(eng.run(train_mcs + 1, rest, eq, coa, 0, 0.0, t.exchange_interval, 0))
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
