"""Learned parallel-tempering ladder.

The reference's ladder is linear in T (src/ether.f90:79-84).  Here a small MLP
surrogate of <E>(T) and log C_v(T), trained by the library's own HIP kernels
on measurements taken with the linear ladder, places the temperatures so that
the thermodynamic length  int sqrt(C_v)/T dT  is equal between neighbours.
For Gaussian energy distributions the swap acceptance between neighbours is
erfc(dT sqrt(C_v) / (2T)), so equal length means uniform acceptance, which
minimises the round-trip time of the deterministic even/odd swap schedule
the reference uses (swap_count parity, src/parallel_tempering.f90:50-62).

The ladder only chooses temperatures: every swap still passes the exact
acceptance test, so the sampled distribution at each temperature is the same
as with the linear ladder.
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
    temps: np.ndarray          # (n,) hottest first
    wall_s: float              # wall time of the production MCS loop
    mcs: int
    round_trips: int           # completed cold->hot->cold trips, all replicas
    acc_rate: np.ndarray       # (n-1,) swap acceptance between slot k and k+1
    updates: float

    @property
    def round_trips_per_s(self) -> float:
        return self.round_trips / self.wall_s if self.wall_s > 0 else 0.0

    @property
    def updates_per_s(self) -> float:
        return self.updates / self.wall_s if self.wall_s > 0 else 0.0


def _ladder_betas(model: Model, temps, repeat):
    return np.tile(model.betas(temps), repeat)


def measure_on_ladder(eng: E.Engine, model: Model, temps, repeat: int, mcs: int, coa: int,
                      pt_interval: int, eq_frac: float = 0.5):
    """Run tempering on `temps` for `mcs` MCS and return per-temperature
    <e> and C_v (the reference's estimators, averaged over repeats)."""
    n = len(temps)
    eng.set_betas(_ladder_betas(model, temps, repeat))
    eng.set_ladders([k * n for k in range(repeat + 1)])
    eng.init_spins(E.INIT_RANDOM)
    eng.reset_moments()
    eq = int(mcs * eq_frac)
    eng.run(1, mcs, eq, coa, 0, 0.0, pt_interval)
    mom = eng.moments()
    beta = model.betas(temps)
    e = np.zeros(n)
    cv = np.zeros(n)
    for k in range(repeat):
        for j in range(n):
            m = mom[k * n + j]
            cnt = m[E.MOM_N]
            ea, e2 = m[E.MOM_E] / cnt, m[E.MOM_E2] / cnt
            e[j] += ea / repeat
            cv[j] += beta[j] ** 2 * (e2 - ea * ea) * model.n_sites / repeat
    return e, cv


def learn_ladder(eng: E.Engine, temps, e, cv, n_out: Optional[int] = None, iters: int = 4000,
                 lr: float = 3e-3, seed: int = 1992):
    """Train the device MLP on (T, e, log C_v) and place a new ladder."""
    temps = np.asarray(temps, np.float64)
    n_out = n_out or len(temps)
    t_lo, t_hi = float(temps.min()), float(temps.max())
    mid, half = 0.5 * (t_lo + t_hi), 0.5 * (t_hi - t_lo)
    x = ((temps - mid) / half).astype(np.float32)
    lcv = np.log(np.maximum(cv, 1e-12))
    y = np.stack([(e - e.mean()) / (e.std() + 1e-30), (lcv - lcv.mean()) / (lcv.std() + 1e-30)],
                 axis=1).astype(np.float32)
    w = np.ones_like(y)
    eng.mlp_init(seed)
    loss = eng.mlp_train(x, y, w, iters, lr)
    T = eng.ladder_place(mid, half, float(lcv.std() + 1e-30), t_lo, t_hi, n_out)
    return T, loss


def run_ladder(eng: E.Engine, model: Model, temps, repeat: int, mcs: int, coa: int,
               pt_interval: int) -> LadderRun:
    """Production tempering run on a given ladder; counts round trips."""
    n = len(temps)
    eng.set_betas(_ladder_betas(model, temps, repeat))
    eng.set_ladders([k * n for k in range(repeat + 1)])
    eng.init_spins(E.INIT_RANDOM)
    eng.reset_moments()
    eng.synchronize()
    t0 = time.perf_counter()
    eng.run(1, mcs, mcs // 4, coa, 0, 0.0, pt_interval)
    eng.synchronize()
    wall = time.perf_counter() - t0
    slot_of, acc, att, trips = eng.pt_state()
    rate = np.zeros(n - 1)
    for k in range(n - 1):
        a = sum(acc[g * n + k] for g in range(repeat))
        b = sum(att[g * n + k] for g in range(repeat))
        rate[k] = a / b if b else 0.0
    return LadderRun(np.asarray(temps), wall, mcs, int(trips.sum()), rate,
                     float(eng.R) * model.n_sites * mcs)


def compare_ladders(model: Model, *, train_mcs: int = 2000, mcs: Optional[int] = None,
                    repeat: Optional[int] = None, iters: int = 4000):
    """Linear ladder versus the learned ladder on the same model (the tempering
    workload of the benchmark).  Returns (linear LadderRun, learned LadderRun, loss)."""
    t = model.tags
    temps = t.temperatures()
    rep = t.repeat if repeat is None else repeat
    mcs = t.tmcs if mcs is None else mcs
    coa = fortran_nint(t.coa)
    eng = make_engine(model, len(temps) * rep)
    try:
        lin = run_ladder(eng, model, temps, rep, mcs, coa, t.exchange_interval)
        e, cv = measure_on_ladder(eng, model, temps, rep, train_mcs, coa, t.exchange_interval)
        T_new, loss = learn_ladder(eng, temps, e, cv)
        learned = run_ladder(eng, model, T_new, rep, mcs, coa, t.exchange_interval)
    finally:
        eng.close()
    return lin, learned, loss
