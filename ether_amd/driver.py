"""Batched ETHER run on the GPU engine: the schedule of src/ether.f90:79-236
with every (temperature, repeat) chain resident on the device at once.

The reference farms temperatures out to MPI ranks and runs REPEAT repeats one
after another per temperature (src/ether.f90:93-214).  Here replica
r = k*nscan + t is repeat k of temperature t; all of them advance together in
one ether_run call per chunk of MCS.  With PT on, each repeat's nscan slots
form one ladder (the reference's ladder is the P ranks of a round,
src/ether.f90:110-133; with one rank per temperature it is the whole list).
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import engine as E
from . import outputs
from .inputs import EtherInputError, GpuTags, Model, fortran_nint


def attach_neighbours(model: Model) -> Model:
    """getting_nbd (src/getting_nbd.f90:46-82) through the library's device
    builder, with the periodic colouring of the lattice; without a GPU the
    host builder and the host colouring give the same lists and colours."""
    t = model.tags
    args = (model.structure.abc, model.basis_xyz, model.sc, [b == "o" for b in t.bc], model.shells, t.nbdfc)
    try:
        cnt, idx, col = E.build_neighbours_device(*args, t.nbd_capacity)
    except E.EtherError as e:
        if e.code != E.ETHER_ERR_NODEV:
            raise
        cnt, idx = E.build_neighbours(*args, t.nbd_capacity)
        col = E.colour_lattice(*args)
    model.nn_count, model.nn_idx, model.colour = cnt, idx, col
    return model


def spiral_partners(model: Model):
    """get_spiral_state('ions_for_spiral') (src/get_spiral_state.f90:47-82) through
    the library's builder: (partners (N, spiral_capacity), spiral_count)."""
    t = model.tags
    return E.build_spiral_partners(model.positions, t.ss_direc, t.ss_dis, t.ss_lat, t.spiral_capacity)


def make_engine(model: Model, n_replicas: int, *, compat: int = 0, seed: Optional[int] = None,
                colour=None) -> E.Engine:
    if model.nn_count is None:
        attach_neighbours(model)
    t = model.tags
    # the spiral state is a vector-spin observable (src/ether.f90:69, src/write_output_files.f90:92)
    spiral = spiral_partners(model)[0] if (t.ssp and not t.ising) else None
    eng = E.Engine(model.model_id, model.n_sites, n_replicas, t.seed if seed is None else seed)
    eng.setup(model.n_species, model.species, model.stagger, model.nn_count, model.nn_idx,
              model.j_exc.reshape(-1), model.spin,
              zeeman=model.h if t.zeeman else None, g_factor=model.g_factor, mu_b=model.mu_b,
              sia=model.sia_vec.reshape(-1) if model.sia_vec is not None else None,
              angle=t.angle, dphi=t.dphi, compat=compat,
              colour=colour if colour is not None else model.colour, spiral=spiral)
    return eng


def apply_gpu_tags(eng: E.Engine, g: GpuTags) -> None:
    """gpu.ether settings that act on a committed context: graph replay,
    launch shape and lanes per site (none of them changes any result)."""
    eng.set_graphs(g.graphs)
    if g.spw:
        eng.set_launch(g.spw, g.xcd)
    if g.lanes:
        eng.set_lanes(g.lanes)


def repeat_observables(mom: np.ndarray, beta: float, n_sites: int) -> np.ndarray:
    """get_eng/get_mag 'avg_*' (src/get_eng.f90:43-56, src/get_mag.f90:44-55).
    Returns (e_avg, U_e, cv, m_avg, U_m, chi)."""
    n = mom[E.MOM_N]
    e, e2, e4 = mom[E.MOM_E] / n, mom[E.MOM_E2] / n, mom[E.MOM_E4] / n
    m, m2, m4 = mom[E.MOM_M] / n, mom[E.MOM_M2] / n, mom[E.MOM_M4] / n
    return np.array([
        e, 1.0 - (1.0 / 3.0) * (e4 / (e2 * e2)), beta * beta * (e2 - e * e) * n_sites,
        m, 1.0 - (1.0 / 3.0) * (m4 / (m2 * m2)), beta * (m2 - m * m) * n_sites,
    ])


def spiral_observables(sums: np.ndarray, n: float, beta: float, count: int) -> np.ndarray:
    """get_spiral_state 'avg_spiral_state' (src/get_spiral_state.f90:122-140).
    Returns (psi_avg, U_psi, chi_psi)."""
    p, p2, p4 = sums[0] / n, sums[1] / n, sums[2] / n
    return np.array([p, 1.5 - 0.5 * (p4 / (p2 * p2)), beta * (p2 - p * p) * count])


def fortran_heads(labels, gaps) -> str:
    """A header record written as the reference writes its character(len=20)
    labels with A13 edits: each label's first 13 bytes, blank-padded on the
    right, after `gaps[k]` blanks of nX editing (src/init.f90:82,
    src/write_output_files.f90:103-104)."""
    out = b""
    for lab, g in zip(labels, gaps):
        out += b" " * g + (lab.encode() + b" " * 20)[:13]
    return out.decode(errors="replace")


FMT103 = (0, 0, 0, 1, 1, 2, 0)
FMT104 = (0, 0, 0, 0, 1, 1, 0, 0)
M_HEAD = ("# 1 Temp.", "2 |M|", "3 χ", "4 Δ|M|", "5 Δχ", "6 U(|M|)", "7 ΔU(|M|)")
SS_HEAD = ("# 1 Temp.", "2 |ψ|", "3 ψ", "4 Δ|ψ|", "5 Δψ", "6 U(|ψ|)", "7 ΔU(|ψ|)")


def e_head(para: bool):
    return ("# 1 Temp.", "2 E", "3 Cv" if para else "3 Cv/kB", "4 ΔE",
            "5 ΔCv" if para else "5 Δ(Cv/kB)", "6 U(E)", "7 ΔU(E)", "8 Accpt. (%)")


def over_repeats(per: np.ndarray):
    """src/observables_per_repeat.f90:29-69: mean and sqrt(sum (x-mean)^2/(m(m-1)))."""
    m = per.shape[0]
    mean = per.mean(axis=0)
    sem = np.sqrt(((per - mean) ** 2).sum(axis=0) / (m * (m - 1.0))) if m > 1 else np.zeros_like(mean)
    return mean, sem


@dataclass
class Results:
    temps: np.ndarray        # (nscan,)
    mean: np.ndarray         # (nscan, 6): e, U_e, cv, m, U_m, chi
    sem: np.ndarray          # (nscan, 6)
    accept_pct: np.ndarray   # (nscan,)
    per_repeat: np.ndarray   # (repeat, nscan, 6)
    wall_s: float
    updates: float
    pt_acc: Optional[np.ndarray] = None
    pt_att: Optional[np.ndarray] = None
    trips: Optional[np.ndarray] = None
    para: bool = False
    spiral_mean: Optional[np.ndarray] = None   # (nscan, 3): psi, U_psi, chi_psi
    spiral_sem: Optional[np.ndarray] = None
    final_S: Optional[np.ndarray] = None       # (nscan, N, 3) at each T after the last repeat
    final_phi: Optional[np.ndarray] = None     # (nscan, N)
    model: Optional[Model] = None
    ladder_source: str = "linear"              # linear, or which trained ladder ran production
    counted_mcs: int = 0                       # MCS whose acceptances accept_pct averages

    def write(self, outdir: str) -> None:
        """energy.dat / magnetization.dat / spiral_state_parameters.dat in the
        reference's layout (src/write_output_files.f90:27-98, rows in format
        (f11.5,6(1x,es12.5),1x,f7.2)).  spiral_state_parameters.dat is created
        in every run and holds rows only with the spiral state on.  With the
        final configurations at hand, ETHER.obs, ETHER.spn, gss.dat,
        moment_vectors.dat and the spK files follow (ether_amd.outputs)."""
        os.makedirs(outdir, exist_ok=True)

        def row(vals, acc=None):
            s = f"{vals[0]:11.5f}" + "".join(f" {v:12.5E}" for v in vals[1:])
            return s + (f" {acc:7.2f}" if acc is not None else "")

        with open(os.path.join(outdir, "magnetization.dat"), "w") as f:
            f.write(fortran_heads(M_HEAD, FMT103) + "\n")
            for k, T in enumerate(self.temps):
                m, s = self.mean[k], self.sem[k]
                f.write(row([T, m[3], m[5], s[3], s[5], m[4], s[4]]) + "\n")
        with open(os.path.join(outdir, "energy.dat"), "w") as f:
            f.write(fortran_heads(e_head(self.para), FMT104) + "\n")
            for k, T in enumerate(self.temps):
                m, s = self.mean[k], self.sem[k]
                f.write(row([T, m[0], m[2], s[0], s[2], m[1], s[1]], self.accept_pct[k]) + "\n")
        with open(os.path.join(outdir, "spiral_state_parameters.dat"), "w") as f:
            if self.spiral_mean is not None:
                f.write(fortran_heads(SS_HEAD, FMT103) + "\n")
                for k, T in enumerate(self.temps):
                    m, s = self.spiral_mean[k], self.spiral_sem[k]
                    f.write(row([T, m[0], m[2], s[0], s[2], m[1], s[1]]) + "\n")
        if self.final_S is not None and self.model is not None:
            outputs.write_all(outdir, self, self.model)


def _schedule(model: Model):
    """(coa, OVRR interval, OVRR passes, PT interval) of src/ether.f90:159-198."""
    t = model.tags
    coa = fortran_nint(t.coa)
    if abs(coa - t.coa) > 1e-12 or coa <= 0:
        raise ValueError("coa must be a positive integer")
    # OVRR gate of src/ether.f90:163-167
    ovrr_on = t.ovrr and t.xyz and not t.zeeman and not t.sia and t.ovrr_mcs > 0
    return (coa, t.ovrr_mcs if ovrr_on else 0, t.ovrr_para if ovrr_on else 0.0,
            t.exchange_interval if t.pt else 0)


def _prepare(model: Model, temps, rep: int, eng: E.Engine) -> None:
    """Temperatures, ladders and the start state (src/ether.f90:139-156)."""
    eng.set_betas(np.tile(model.betas(temps), rep))
    if model.tags.pt:
        eng.set_ladders([k * len(temps) for k in range(rep + 1)])
    # start state: generate_supercell state, then fresh_spins (src/ether.f90:155)
    eng.set_spins(-1, model.initial_spins, model.initial_phi)
    eng.init_spins(E.INIT_FRESH)
    eng.reset_moments()


def _train(model: Model, temps, rep: int, eng: E.Engine, sched):
    """gpu.ether `ladder = learned`: train the ladder inside the equilibration
    MCS and keep the one with the best measured round-trip cost.  Returns
    (temperatures, next step, swap count, source)."""
    t, g = model.tags, model.gpu
    if not t.pt:
        raise EtherInputError("ladder = learned needs parallel tempering (pt = .T.)")
    ntr = g.train_mcs or t.tmcs_eq // 2
    if ntr <= 0 or ntr + 3 * (ntr // 8) > t.tmcs_eq:
        raise EtherInputError("ladder = learned: the training MCS (and three validation windows "
                              "of 1/8 of them) must lie within the equilibration MCS")
    from .ladder import train_ladder   # ladder imports this module
    coa, ovrr_iv, ovrr_p, pt_iv = sched
    tr = train_ladder(eng, model, temps, rep, coa, pt_iv, ntr, windows=g.windows,
                      ovrr_iv=ovrr_iv, ovrr_p=ovrr_p)
    eng.set_ladders([k * len(temps) for k in range(rep + 1)])
    eng.reset_moments()
    return np.asarray(tr.ladder), tr.next_step, tr.swap_count, tr.source


def _results(model: Model, eng: E.Engine, temps, rep: int, mom: np.ndarray, wall: float,
             counted: int, source: str, spiral_sums: Optional[np.ndarray] = None) -> Results:
    t = model.tags
    nscan = len(temps)
    beta_T = model.betas(temps)
    per = np.zeros((rep, nscan, 6))
    acc = np.zeros(nscan)
    for k in range(rep):
        for j in range(nscan):
            s = k * nscan + j
            per[k, j] = repeat_observables(mom[s], beta_T[j], model.n_sites)
            acc[j] += mom[s, E.MOM_ACC]
    mean = np.zeros((nscan, 6))
    sem = np.zeros((nscan, 6))
    for j in range(nscan):
        mean[j], sem[j] = over_repeats(per[:, j, :])
    res = Results(temps=np.asarray(temps), mean=mean, sem=sem,
                  accept_pct=acc / (rep * counted) * 100.0, per_repeat=per, wall_s=wall,
                  updates=float(eng.R) * model.n_sites * counted, para=t.para,
                  ladder_source=source, counted_mcs=counted)
    if getattr(eng, "spiral_count", 0):
        sums = spiral_sums if spiral_sums is not None else eng.spiral()[0]
        sper = np.zeros((rep, nscan, 3))
        for k in range(rep):
            for j in range(nscan):
                s = k * nscan + j
                sper[k, j] = spiral_observables(sums[s], mom[s, E.MOM_N], beta_T[j], eng.spiral_count)
        res.spiral_mean = np.zeros((nscan, 3))
        res.spiral_sem = np.zeros((nscan, 3))
        for j in range(nscan):
            res.spiral_mean[j], res.spiral_sem[j] = over_repeats(sper[:, j, :])
    if t.pt:
        _, pacc, patt, trips = eng.pt_state()
        res.pt_acc, res.pt_att, res.trips = pacc, patt, trips
    # the configuration at each temperature after the last repeat: the last
    # ladder of the context (all of it when the repeats ran one after another)
    res.final_S, res.final_phi = outputs.final_states(eng, nscan, eng.R // nscan)
    res.model = model
    return res


def run(model: Model, *, compat: int = 0, chunk: int = 0, temps=None, repeat: Optional[int] = None,
        eng: Optional[E.Engine] = None, seed: Optional[int] = None) -> Results:
    """Full run: every temperature x repeat as one batch of device replicas.
    The run directory's gpu.ether settings (model.gpu) apply on top."""
    t, g = model.tags, model.gpu
    temps = t.temperatures() if temps is None else np.asarray(temps, np.float64)
    nscan = len(temps)
    rep = t.repeat if repeat is None else repeat
    flags = compat | g.compat
    chain = bool(flags & E.COMPAT_CHAIN_REPEATS)
    if chain and g.ladder == "learned":
        raise EtherInputError("ladder = learned does not run with chained repeats (compat 8)")
    # every repeat its own replica, or (chained) the repeats one after another on nscan replicas
    R = nscan if chain else nscan * rep
    own = eng is None
    if own:
        eng = make_engine(model, R, compat=flags, seed=seed)
        apply_gpu_tags(eng, g)
    sched = _schedule(model)
    coa, ovrr_iv, ovrr_p, pt_iv = sched
    _prepare(model, temps, R // nscan, eng)
    if g.autotune:
        eng.autotune(g.autotune)      # restores every replica's state
    t0 = time.perf_counter()
    step, swap, source = 1, 0, "linear"
    if g.ladder == "learned":
        temps, step, swap, source = _train(model, temps, rep, eng, sched)
    first = step
    chunk = g.chunk or chunk or t.tmcs
    moms, sums = [], []
    for k in range(rep if chain else 1):
        if k > 0:
            # the next repeat: fresh_spins of the state the last one left, zeroes('avg')
            # (src/ether.f90:152-156); swap_count restarts (:158)
            eng.init_spins(E.INIT_FRESH)
            eng.reset_moments()
            step, swap = 1, 0
        while step <= t.tmcs:
            n = min(chunk, t.tmcs - step + 1)
            swap = eng.run(step, n, t.tmcs_eq, coa, ovrr_iv, ovrr_p, pt_iv, swap)
            step += n
        moms.append(eng.moments())   # synchronises
        if getattr(eng, "spiral_count", 0):
            sums.append(eng.spiral()[0])
    mom = np.concatenate(moms)
    wall = time.perf_counter() - t0
    res = _results(model, eng, temps, rep, mom, wall, t.tmcs - first + 1, source,
                   np.concatenate(sums) if sums else None)
    if own:
        eng.close()
    return res


def _build_engines(jobs, engs: list, workers: int = 8):
    """Committed, prepared contexts for (model, seed) pairs, set up on several
    host threads at once (each context's set-up is independent and runs in
    native code).  Every context built is appended to `engs`, in job order,
    so that the caller closes them all when one of them fails; the first
    failure is raised after every job has ended."""
    from concurrent.futures import ThreadPoolExecutor

    def one(job):
        m, seed = job
        tk = m.tags.temperatures()
        e = make_engine(m, len(tk) * m.tags.repeat, compat=m.gpu.compat, seed=seed)
        try:
            apply_gpu_tags(e, m.gpu)
            _prepare(m, tk, m.tags.repeat, e)
        except BaseException:
            e.close()
            raise
        return e, tk

    E.load_library()
    with ThreadPoolExecutor(max_workers=max(1, min(workers, len(jobs)))) as pool:
        futs = [pool.submit(one, j) for j in jobs]
        done = [f.exception() is None for f in futs]   # waits for every job
    built = [f.result() for f, ok in zip(futs, done) if ok]
    engs.extend(e for e, _ in built)
    for f, ok in zip(futs, done):
        if not ok:
            raise f.exception()
    return built


def run_batch(models: Sequence[Model], *, seeds: Optional[Sequence[Optional[int]]] = None) -> List[Results]:
    """A batch manifest's runs (batch.ether): each run directory is its own
    context (lattice, couplings, temperatures, repeats, seed); runs whose
    schedules are equal (mcs, coa, OVRR and PT tags) and that share the spin
    model and the SIA switch advance together through one batched schedule
    (ether_batch_run), one launch per kernel kind for the whole group.  Every
    run's results equal those of run() on it alone."""
    seeds = list(seeds) if seeds is not None else [None] * len(models)
    for m in models:
        if m.gpu.ladder != "linear" or m.gpu.autotune or m.gpu.compat & E.COMPAT_CHAIN_REPEATS:
            raise EtherInputError("a batched run takes neither `ladder = learned`, `autotune` nor "
                                  "chained repeats (compat 8)")
    groups = {}
    for k, m in enumerate(models):
        key = (m.tags.tmcs, m.tags.tmcs_eq) + _schedule(m) + (m.model_id, m.sia_vec is not None)
        groups.setdefault(key, []).append(k)
    out: List[Optional[Results]] = [None] * len(models)
    for key, idx in groups.items():
        tmcs, eq, coa, ovrr_iv, ovrr_p, pt_iv = key[:6]
        engs, temps, batch = [], [], None
        try:
            for e, tk in _build_engines([(models[k], seeds[k]) for k in idx], engs):
                temps.append(tk)
            batch = E.Batch(engs)
            batch.set_graphs(all(models[k].gpu.graphs for k in idx))
            t0 = time.perf_counter()
            chunk = min(models[k].gpu.chunk or tmcs for k in idx)
            step, swap = 1, 0
            while step <= tmcs:
                n = min(chunk, tmcs - step + 1)
                swap = batch.run(step, n, eq, coa, ovrr_iv, ovrr_p, pt_iv, swap)
                step += n
            moms = [e.moments() for e in engs]
            wall = time.perf_counter() - t0
            for j, k in enumerate(idx):
                out[k] = _results(models[k], engs[j], temps[j], models[k].tags.repeat, moms[j], wall,
                                  tmcs, "linear")
        finally:
            if batch is not None:
                batch.close()          # before its contexts
            for e in engs:
                e.close()
    return out
