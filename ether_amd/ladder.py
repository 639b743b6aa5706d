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
Gaussian formula:  sqrt(C(T_mid)) = 2 erfcinv(a_k) T_mid / dT_k.  The
surrogate is trained on that acceptance-consistent C (and on <e>), measured on
the linear ladder during the first train_mcs MCS of the run.

Protocol (adaptive_run): the first train_mcs MCS of the equilibration run are
split into windows.  The first runs on the linear ladder; after each window
the MLP is retrained on the samples of all windows so far and the temperatures
of the slots are retuned in place (every replica keeps its state and slot), so
a pair whose acceptance was too small to measure on one ladder is measured on
the next.  The remaining MCS run on the last ladder.  Every swap passes the
exact acceptance test throughout, so the sampled distribution at each
temperature is unchanged; only the production window after the last retune
is timed and its round trips counted.
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


def acceptance_heat_capacity(temps, acc, attempts):
    """Heat capacity that reproduces the measured pair acceptance through
    a = erfc(dT sqrt(C) / (2 T)), at the pair midpoints, interpolated to temps."""
    temps = np.asarray(temps, np.float64)
    a = np.clip(np.asarray(acc, np.float64), 0.5 / np.maximum(attempts, 1.0), 0.999)
    x = 2.0 * erfcinv(a)
    tm = 0.5 * (temps[1:] + temps[:-1])
    dt = np.abs(np.diff(temps))
    c = (x * tm / dt) ** 2
    order = np.argsort(tm)
    return np.interp(temps, tm[order], c[order])


def learn_ladder(eng: E.Engine, temps, e, c, n_out: Optional[int] = None, iters: int = 4000,
                 lr: float = 3e-3, seed: int = 1992, weight=None, t_range=None):
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
    w = np.ones_like(y) if weight is None else np.repeat(
        np.asarray(weight, np.float32)[:, None], 2, axis=1)
    eng.mlp_init(seed)
    loss = eng.mlp_train(x, y, w, iters, lr)
    T = eng.ladder_place(mid, half, float(lc.std() + 1e-30), t_lo, t_hi, n_out)
    return T, loss


def _window(eng: E.Engine, model: Model, temps, repeat: int, first: int, n_mcs: int, coa: int,
            pt_interval: int, sc: int):
    """Run n_mcs MCS from step `first` with fresh swap counters and energy
    moments; returns (pair acceptance, attempts, <e> per slot, swap counter)."""
    n = len(temps)
    eng.set_ladders([k * n for k in range(repeat + 1)])
    eng.reset_moments()
    # every step of the window samples (eq = first - 1) so that <e> is measured
    sc = eng.run(first, n_mcs, first - 1, coa, 0, 0.0, pt_interval, sc)
    mom = eng.moments()
    a, att, _ = _pair_acceptance(eng, n, repeat)
    e = np.zeros(n)
    for g in range(repeat):
        for k in range(n):
            m = mom[g * n + k]
            e[k] += m[E.MOM_E] / max(m[E.MOM_N], 1.0) / repeat
    return a, att, e, sc


def adaptive_run(eng: E.Engine, model: Model, temps, repeat: int, mcs: int, coa: int,
                 pt_interval: int, *, learn: bool, train_mcs: int = 2000,
                 windows: int = 4) -> LadderRun:
    """Tempering run of mcs MCS.  The first train_mcs MCS are `windows` equal
    windows; each one measures the pair acceptance and <e> of the ladder it ran
    on.  learn=True: after every window the MLP is retrained on the samples of
    all windows so far (window k weighted 2^k, later ladders are closer to the
    final one) and the slots are retuned to the ladder it places, so each
    window refines the acceptance-consistent C where the previous ladder was
    too coarse to measure it.  learn=False: every window runs on `temps`.
    The remaining mcs - train_mcs MCS are the timed production window."""
    temps = np.asarray(temps, np.float64)
    n = len(temps)
    t_range = (float(temps.min()), float(temps.max()))
    eq = model.tags.tmcs_eq
    eng.set_betas(np.tile(model.betas(temps), repeat))
    eng.init_spins(E.INIT_RANDOM)
    win = max(train_mcs // windows, 1)
    cur, sc, first = temps, 0, 1
    xs, es, cs, ws = [], [], [], []
    a = c = e = None
    loss = None
    for k in range(windows):
        length = win if k < windows - 1 else train_mcs - win * (windows - 1)
        a, att, e, sc = _window(eng, model, cur, repeat, first, length, coa, pt_interval, sc)
        first += length
        c = acceptance_heat_capacity(cur, a, att)
        xs.append(cur)
        es.append(e)
        cs.append(c)
        ws.append(np.full(n, 2.0 ** k))
        if learn:
            cur, loss = learn_ladder(eng, np.concatenate(xs), np.concatenate(es),
                                     np.concatenate(cs), n_out=n, weight=np.concatenate(ws),
                                     t_range=t_range)
            eng.retune_betas(np.tile(model.betas(cur), repeat))
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
    return LadderRun(cur, wall, rest, trips, acc, float(eng.R) * model.n_sites * rest,
                     train_e=e, train_c=c, train_acc=a, mlp_loss=loss)


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
