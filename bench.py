#!/usr/bin/env python3
"""Benchmark: spin updates/s of the Monte Carlo sweep engine on one MI355X.

Workload (default): Heisenberg antiferromagnet on a 32^3 simple-cubic lattice
(32768 FP32 unit vectors), J = 1 meV, Marsaglia moves, overrelaxation every
MCS, 19 temperatures 60 -> 6 K x REPEAT 10 = 190 independent chains, all
resident on the GPU and advanced together; MCS 5000 with 3000 equilibration,
observables every 10 MCS.  A "step" is one MCS of every chain (OVRR +
Metropolis sweep + the scheduled measurement), the per-MCS schedule of the
reference's driver, with the reflection fused into each colour phase of the
sweep and the measurement booked by the sweep before it (two colours).  The
launch shape of the sweep is picked by the autotuner (`--launch` pins it) on
the start state after fresh_spins, which it restores; `--warmup` untimed
production MCS follow; the timed region is `--steps` production MCS (samples
every 10 MCS included) from the next schedule period, whole periods replayed
as HIP graphs that are captured and uploaded before the warmup.

value = Metropolis spin updates / wall second of the timed MCS loop
        (OVRR and measurement time included, not counted as updates).
roofline = the Metropolis colour-phase kernel: 24 B of own-spin traffic per
        update (FP32 SoA read + write) / its average launch time from HIP
        events on the engine stream (one event pair around the colour phases
        of each sweep over `--event-steps` more MCS of the same run, launched
        one by one), against 8 TB/s; traffic = HBM bytes per launch at the
        shape the run keeps, from the committed PMC summary under profiles/.
pt, pt_trips = the tempering ladder comparisons (linear vs learned).
cpu_baseline = the CPU oracle's FP64 sweep (+OVRR) of 32^3 chains, one per
        host thread (up to 16) after one alone on one core, for a bounded
        sample of about 10 s.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin updates/s; PT round-trips/s (pyrochlore, learned vs plain ladder)"
B_UPD = {0: 2.0, 1: 24.0}   # algorithmic bytes per update: Ising int8, XYZ FP32 SoA
HBM_PEAK_GBS = 8000.0


def cpu_baseline(model, budget_s=10.0, threads=None):
    """Oracle sweeps (FP64 C, an OVRR pass then a Metropolis sweep per MCS, as
    the headline schedules them) on the host cores: one chain per thread, as
    the reference runs one temperature per MPI rank (src/ether.f90:110-133),
    after one chain alone on one core for a quarter of the budget.  The oracle
    keeps no global state and ctypes releases the GIL around each call."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from oracle import oracle as o
    from ether_amd.driver import attach_neighbours
    if model.nn_count is None:
        attach_neighbours(model)
    orc = o.Oracle.from_model(model)
    n = model.n_sites
    # two-colour order of the simple-cubic lattice (what the GPU runs)
    L = model.sc
    idx = np.arange(n)
    x, y, z = idx % L[0], (idx // L[0]) % L[1], idx // (L[0] * L[1])
    order = np.concatenate([idx[(x + y + z) % 2 == 0], idx[(x + y + z) % 2 == 1]]).astype(np.int32)
    rid = np.arange(n, dtype=np.int32)
    beta = model.betas([20.0])[0]

    def chain(k, seconds):
        v = np.random.default_rng(1 + k).normal(size=(n, 3))
        S = np.ascontiguousarray(v / np.linalg.norm(v, axis=1)[:, None])
        t0 = time.perf_counter()
        sweeps = 0
        while time.perf_counter() - t0 < seconds:
            orc.overrelax_pass(S, None, order, rid, 1.0, 1992, k, sweeps, 0)
            orc.sweep(S, None, order, rid, beta, 1992, k, sweeps)
            sweeps += 1
        return sweeps, time.perf_counter() - t0

    s1, dt1 = chain(0, budget_s / 4)
    one = s1 * n / dt1
    # the host's share of cores for one GPU: at most 16 threads
    T = threads or max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=T) as pool:
        res = list(pool.map(lambda k: chain(k, budget_s * 3 / 4), range(T)))
    wall = time.perf_counter() - t0
    total = sum(sw for sw, _ in res)
    return {"value": total * n / wall, "unit": "spin updates/s", "cores": T, "kind": "port",
            "single_core_value": one,
            "sample": f"{T} chains of a {L[0]}x{L[1]}x{L[2]} Heisenberg lattice at T=20 K on {T} threads, "
                      f"{total} MCS (OVRR pass + Metropolis sweep) in {wall:.1f} s, FP64 C oracle; one "
                      f"chain alone on one core: {s1} MCS in {dt1:.1f} s"}


def describe(name, m, n_temps, R, steps, split_ovrr, coa, ovrr_iv, measure_pass=False):
    t = m.tags
    kind = "Ising" if t.ising else "Heisenberg"
    moves = "" if t.ising else (f", ANGLE moves delta {math.degrees(t.dphi):g} deg" if t.angle else ", Marsaglia moves")
    extra = (", Zeeman" if t.zeeman else "") + (", SIA" if t.sia else "")
    ovrr = ""
    if ovrr_iv:
        ovrr = (f", OVRR every {ovrr_iv} MCS "
                f"({'split pass' if split_ovrr else 'fused into the sweep phases'})")
    return (f"{name}: {kind} {m.sc[0]}x{m.sc[1]}x{m.sc[2]} cells ({m.n_sites} spins){extra}, "
            f"{n_temps} T x {t.repeat} repeats = {R} chains{moves}{ovrr}, "
            f"MCS {steps}/{t.tmcs_eq}, coa {coa}" + (f", swaps every {t.exchange_interval} MCS on one ladder "
                                                     f"per repeat" if t.pt else "") +
            (", measurement as a pass of its own" if measure_pass else ""))


def profiled_traffic(kernel, workload, key):
    """HBM bytes per launch (or per unit) of the dominant kernel from the
    committed PMC summary under profiles/ (FETCH_SIZE x2.0 + WRITE_SIZE x1.0,
    factors from a plain copy): the line `<key> = bytes` of
    profiles/<kernel>_<workload>_traffic.md or _stats.md, or None."""
    for suffix in ("traffic", "stats"):
        path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_{suffix}.md")
        if not os.path.exists(path):
            continue
        with open(path) as f:
            for line in f:
                if line.startswith(key + " ="):
                    return float(line.split("=")[-1])
    return None


def profiled_value(kernel, workload, suffix, key):
    """The line `<key> = value` of profiles/<kernel>_<workload>_<suffix>.md, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_{suffix}.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + " ="):
                return float(line.split("=")[-1])
    return None


def valu_roofline(kernel, workload, spw, xcd, resident, avg_ms):
    """VALU utilisation next to the HBM fraction: the committed PMC summary's
    VALU issue cycles per SIMD (SQ_ACTIVE_INST_VALU x 4 / SIMDs) over the
    launch time this run measured, at 2.4 GHz."""
    if resident:
        cyc = profiled_value(kernel, workload, "valu", "valu_cycles_per_simd_per_sweep")
        simds = profiled_value(kernel, workload, "valu", "valu_simds_used")
    else:
        cyc = profiled_value(kernel, workload, "valu", f"valu_cycles_per_simd[spw={spw},xcd={xcd}]")
        simds = 1024.0
    if cyc is None or not avg_ms:
        return None
    floor_ms = cyc / 2.4e6
    return {"issue_floor_ms": floor_ms, "busy": floor_ms / avg_ms, "simds": simds,
            "basis": "VALU issue cycles per SIMD from the committed PMC summary "
                     f"(profiles/{kernel}_{workload}_valu.md) over this run's launch time"
                     + (" (per sweep, on the SIMDs of the CUs the launch occupies)" if resident else "")}


def copy_probe(n_bytes, reps):
    """A plain device-to-device copy of known size (the HIP runtime's copy
    kernel), used to calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n_bytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n_bytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n_bytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n_bytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n_bytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run only a device copy of this many bytes (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="sc_heis_afm_32")
    ap.add_argument("--event-steps", type=int, default=500,
                    help="MCS after the timed region whose kernels are timed with HIP events")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pt", action="store_true", help="skip the tempering ladder comparison")
    ap.add_argument("--pt-mcs", type=int, default=20000)
    ap.add_argument("--pt-trips-mcs", type=int, default=200000,
                    help="MCS of the round-trip comparison on 128 spins")
    ap.add_argument("--no-sq", action="store_true", help="skip the structure-factor timing")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven workload")
    ap.add_argument("--batch-mcs", type=int, default=100)
    ap.add_argument("--batch-kinds", default="",
                    help="comma-separated lattice kinds of the batch section (sc_ising, sc_heis, tri, pyro; default all)")
    ap.add_argument("--no-batch-plain", action="store_true",
                    help="skip the batch section's one-launch-per-lattice leg")
    ap.add_argument("--sq-calls", type=int, default=5)
    ap.add_argument("--no-ising", action="store_true", help="skip the Ising simple-cubic section")
    ap.add_argument("--ising-steps", type=int, default=2000)
    ap.add_argument("--tag", action="append", default=[], metavar="KEY=VALUE",
                    help="replace an in.ether tag of the headline workload (e.g. angle='.F. 0')")
    ap.add_argument("--no-c3", action="store_true", help="skip the triangular ANGLE/Marsaglia sweep section")
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--split-ovrr", action="store_true",
                    help="the reference's order: a whole OVRR pass, then the sweep")
    ap.add_argument("--measure-pass", action="store_true",
                    help="every measurement as a pass of its own after the sweep")
    ap.add_argument("--launch", type=int, nargs=2, metavar=("SPW", "XCD"), default=None,
                    help="sites per wave and XCD map of the sweep instead of the autotuner's choice")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe, 4)
        return
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); the replicas "
              "of a run are independent and would shard across GPUs by whole ladders", file=sys.stderr)
        sys.exit(2)

    over = dict(kv.split("=", 1) for kv in args.tag)
    out = measure(args.workload, args.steps, args.warmup, args.event_steps, args.split_ovrr, args,
                  over=over, label=(args.workload + " with " + ", ".join(f"{k} = {v}" for k, v in over.items()))
                  if over else None)
    out = {"metric": METRIC, **out}
    if not args.no_pt:
        out["pt"] = pt_comparison(args.pt_mcs)
        # On 1 -> 0.05 K (T/J 0.086 -> 0.004) uniform Marsaglia proposals are
        # accepted at about kT/4J or less, the chains' energies hardly move
        # between swaps and no label completes a round trip on any ladder; and
        # at 65536 spins the equilibrium acceptance of a 64-temperature ladder,
        # erfc(sqrt(C) dlnT / 2) with C ~ N, is ~1e-17 even for the best
        # ladder.  Round trips are counted on 2^3 cells (128 spins) on
        # 20 -> 1 K with the same 64 temperatures and schedule, over enough
        # swap rounds for labels to cross a 64-slot ladder.
        out["pt_trips"] = pt_comparison(args.pt_trips_mcs, cells=2, temp="20 1 0.301587301587302")
        # the second metric in brief, where a reader of the config sees it
        out["config"]["pt_round_trips"] = {k: pt_brief(out[k]) for k in ("pt", "pt_trips")}
    if not args.no_sq:
        out["sq"] = sq_bench(args.sq_calls)
    if not args.no_batch:
        out["batch"] = batch_bench(args.batch_mcs, kinds=[k for k in args.batch_kinds.split(",") if k],
                                   plain=not args.no_batch_plain)
    if not args.no_ising and args.workload != "sc_ising_16":
        ising = measure("sc_ising_16", args.ising_steps, 20, 100, False, args)
        out["ising"] = {k: ising[k] for k in ("value", "unit", "ms_per_step", "config", "roofline",
                                              "launch", "acceptance_mean")}
    if not args.no_c3 and args.workload != "tri_afm_96":
        out["c3"] = c3_bench(args.c3_steps, args)
    if not args.no_cpu:
        from ether_amd import workloads
        from ether_amd.driver import attach_neighbours
        m = workloads.model(args.workload)
        attach_neighbours(m)
        out["cpu_baseline"] = cpu_baseline(m, args.cpu_budget)
    print(json.dumps(out), flush=True)


def measure(workload, steps, warmup, event_steps, split_ovrr, args, over=None, label=None):
    """The headline measurement on one workload: autotuned launch shape, the
    start state and fresh_spins, `warmup` untimed production MCS, `steps`
    timed production MCS of the schedule from the next period boundary (whole
    periods replayed as graphs captured before the timed region), then
    `event_steps` more MCS launched one by one between HIP events for the
    kernel durations.  `over` replaces in.ether tags of the workload; `label`
    names the variant (profiles and the config line)."""
    import numpy as np
    from ether_amd import engine as E
    from ether_amd import workloads
    from ether_amd.driver import attach_neighbours, make_engine
    from ether_amd.inputs import fortran_nint

    m = workloads.model(workload, **(over or {}))
    workload = label or workload
    attach_neighbours(m)
    t = m.tags
    temps = t.temperatures()
    R = len(temps) * t.repeat
    eng = make_engine(m, R, compat=(E.COMPAT_SPLIT_OVRR if split_ovrr else 0) |
                      (E.COMPAT_MEASURE_PASS if args.measure_pass else 0))
    eng.set_betas(np.tile(m.betas(temps), t.repeat))
    # the start state and fresh_spins
    eng.set_spins(-1, m.initial_spins, m.initial_phi)
    eng.init_spins(E.INIT_FRESH)
    # launch shape: the autotuner times every candidate on this context and
    # puts the spins back; it is the last long GPU work before the timed
    # region, so the GPU starts that region out of its idle power state
    if args.launch:
        spw, xcd = args.launch
        eng.set_launch(spw, xcd)
        tune_ms = None
    else:
        spw, xcd, tune_ms = eng.autotune(reps=40)
    coa = fortran_nint(t.coa)
    ovrr_on = t.ovrr and t.xyz and not t.zeeman and not t.sia
    ovrr_iv = t.ovrr_mcs if ovrr_on else 0
    ovrr_p = t.ovrr_para if ovrr_on else 0.0

    # tempering (pt = .T.): one ladder per repeat, swaps every exchange_interval MCS
    pt_iv = t.exchange_interval if t.pt else 0
    sc = 0
    if pt_iv:
        eng.set_ladders([k * len(temps) for k in range(t.repeat + 1)])
    # the schedule's period: whole periods are replayed as captured graphs
    eq = t.tmcs_eq
    period = coa
    for iv in (ovrr_iv, pt_iv):
        if iv:
            period = period * iv // math.gcd(period, iv)
    # the timed window is K production MCS (samples included) from the period
    # boundary after the warmup; every graph it replays is captured and
    # uploaded before the warmup, so that the warmup runs right before the
    # clock starts
    first = eq + 1 + period * -(-warmup // period)
    eng.run_prepare(first, steps, eq, coa, ovrr_iv, ovrr_p, pt_iv)
    # warmup: W untimed production MCS
    if warmup > 0:
        sc = eng.run(eq + 1, warmup, eq, coa, ovrr_iv, ovrr_p, pt_iv, sc)
    eng.reset_moments()
    eng.synchronize()
    t0 = time.perf_counter()
    sc = eng.run(first, steps, eq, coa, ovrr_iv, ovrr_p, pt_iv, sc)
    eng.synchronize()
    wall = time.perf_counter() - t0
    mom = eng.moments()
    updates = float(R) * m.n_sites * steps
    value = updates / wall

    # event-timed continuation of the same run (the next E steps): every
    # kernel is launched on its own between HIP events on the engine stream,
    # which a graph replay does not allow; gives the per-launch kernel durations
    eng.set_timing(True)
    t1 = time.perf_counter()
    eng.run(first + steps, event_steps, eq, coa, ovrr_iv, ovrr_p, pt_iv, sc)
    eng.synchronize()
    wall_ev = time.perf_counter() - t1
    sweep_ms, sweep_n = eng.timing(E.TIME_SWEEP)
    ovrr_ms, ovrr_n = eng.timing(E.TIME_OVRR)
    meas_ms, meas_n = eng.timing(E.TIME_MEASURE)
    smeas_ms, smeas_n = eng.timing(E.TIME_SWEEP_MEAS)
    res_ms, res_n = eng.timing(E.TIME_RESIDENT)
    eng.set_timing(False)

    nc, col, _ = eng.layout()
    if res_n > 0:
        # Ising with the lattice resident in LDS: one launch runs the sweeps
        # between two samples; the unit is one sweep of every replica (a
        # workgroup per replica, or per word of four replicas)
        kernel = "k_ising_resident1" if eng.resident_kind() == 3 else "k_ising_resident"
        unit = "one sweep (all colour phases) inside a resident launch"
        per_launch_updates = R * m.n_sites
        avg_ms = res_ms / res_n
        sweep_ms, sweep_n = res_ms, res_n
    else:
        kernel = "k_ising_sweep" if m.model_id == 0 else "k_xyz_sweep"
        unit = "one colour phase"
        per_launch_updates = R * m.n_sites / nc
        avg_ms = sweep_ms / max(sweep_n, 1)
    b_upd = B_UPD[m.model_id] + (8.0 if (m.model_id == 1 and t.angle) else 0.0)
    achieved = b_upd * per_launch_updates / (avg_ms * 1e-3) / 1e9
    out = {
        "value": value,
        "unit": "spin updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (random-start spins, reference in.ether tags of the workload)",
        "config": {
            "workload": describe(workload, m, len(temps), R, steps, split_ovrr, coa, ovrr_iv,
                                 args.measure_pass),
            "replicas": R, "sites": m.n_sites, "colours": nc, "mcs": steps,
            "timed_steps": f"production MCS {first}..{first + steps - 1} (equilibration {eq}), "
                           f"after {warmup} untimed production MCS",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(kernel, workload, "traffic_bytes_per_sweep" if res_n > 0 else
                                        f"traffic_bytes_per_launch[spw={spw},xcd={xcd}]"),
            "kernel": f"{kernel} ({unit})",
            "valu": valu_roofline(kernel, workload, spw, xcd, res_n > 0, avg_ms),
            "bytes_per_update": b_upd, "updates_per_launch": per_launch_updates,
            "avg_launch_ms": avg_ms, "launches": sweep_n,
            "timing": (f"HIP events on the engine stream around the back-to-back colour phases of "
                       f"every sweep over MCS {first + steps}..{first + steps + event_steps - 1} of the same run "
                       f"(launched one by one; {wall_ev / event_steps * 1e3:.4f} ms per MCS there), "
                       f"divided by the phases they cover") if res_n == 0 else
                      (f"HIP events on the engine stream around every resident launch over MCS "
                       f"{first + steps}..{first + steps + event_steps - 1} of the same run "
                       f"({wall_ev / event_steps * 1e3:.4f} ms per MCS there), divided by the sweeps they cover"),
        },
        "launch": {"sites_per_wave": spw, "xcd_map": xcd,
                   "autotune_ms_per_sweep": None if tune_ms is None else
                   {f"spw{k}_xcd{x}": round(float(tune_ms[x][i]), 5)
                    for x in range(E.Engine.AUTOTUNE_MAPS) for i, k in enumerate(E.Engine.AUTOTUNE_SPW)}},
        "kernel_ms": {"sweep": sweep_ms, "ovrr": ovrr_ms, "measure": meas_ms,
                      "sweep_measuring": smeas_ms,
                      "sweep_launches": sweep_n, "ovrr_launches": ovrr_n, "measure_launches": meas_n,
                      "sweep_measuring_launches": smeas_n,
                      "note": "sweep: colour phases of the sweeps no measurement follows (the roofline "
                              "kernel); sweep_measuring: those of the sweeps whose phases also book "
                              "the measurement that follows (XYZ); measure: the measurement passes, "
                              "or only their finalize after a sweep that booked them"},
        "acceptance_mean": float(mom[:, E.MOM_ACC].mean() / steps),
        "graphs": "whole periods of P = lcm(coa, OVRR, PT intervals) MCS replayed as captured HIP graphs",
    }
    eng.close()
    return out


def c3_bench(steps, args):
    """Sweep throughput on the triangular workload (96 x 96 cells, SIA and
    Zeeman, 32 T x 10 repeats = 320 chains, OVRR off as in the reference with
    a field or anisotropy) for both move types: ANGLE moves of 5 degrees and
    Marsaglia moves (src/update_spin_details.f90:39-46), each timed as the
    headline is (production MCS, graphs captured before the clock)."""
    keep = ("value", "unit", "ms_per_step", "config", "roofline", "launch", "acceptance_mean")
    res = {}
    for name, over in (("angle", {}), ("marsaglia", {"angle": ".F. 0"})):
        r = measure("tri_afm_96", steps, 20, 100, False, args, over=over,
                    label="tri_afm_96" if name == "angle" else "tri_afm_96_marsaglia")
        res[name] = {k: r[k] for k in keep}
    per_type = 32 * 10 * 5000 * 9216          # spin updates of the whole run per move type
    res["spin_updates_per_move_type"] = per_type
    res["run_s_per_move_type"] = {k: per_type / res[k]["value"] for k in ("angle", "marsaglia")}
    res["angle_over_marsaglia"] = res["angle"]["value"] / res["marsaglia"]["value"]
    return res


def sq_bench(calls):
    """S(q) of every replica of the triangular workload: 96 x 96 cells (9216
    spins, 9216 commensurate q points), 32 T x 10 repeats = 320 replicas,
    after 200 MCS from a random start.  One fused complex DFT GEMM per
    lattice axis on the matrix cores (split-BF16, FP32 accumulation), the
    first reading the spin planes, the last forming |S(q)|^2; then a
    transpose to replica-major output.  The device time comes from
    HIP events around the device part of each call; the error is measured
    against numpy's FP64 FFT for 4 replicas."""
    import numpy as np
    from ether_amd import engine as E
    from ether_amd import workloads
    from ether_amd.driver import make_engine
    m = workloads.model("tri_afm_96")
    t = m.tags
    temps = t.temperatures()
    R = len(temps) * t.repeat
    eng = make_engine(m, R)
    eng.set_betas(np.tile(m.betas(temps), t.repeat))
    eng.init_spins(E.INIT_RANDOM)
    eng.run(1, 200, 10 ** 6, 10)
    L = (m.sc[0], m.sc[1], m.sc[2])
    sq = eng.structure_factor(L)                 # first call allocates and builds the tables
    eng.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(calls):
        sq = eng.structure_factor(L)
    wall = (time.perf_counter() - t0) / calls
    dev_ms, n = eng.timing(E.TIME_SQ)
    eng.set_timing(False)
    worst = 0.0
    for r in range(4):
        S, _ = eng.get_spins(r)
        grid = S.reshape(L[1], L[0], 3)
        ref = sum(np.abs(np.fft.fft2(grid[:, :, a])) ** 2 for a in range(3)).reshape(-1) / m.n_sites ** 2
        worst = max(worst, float(np.abs(sq[r] - ref).max() / ref.max()))
    eng.close()
    N = m.n_sites
    nr3 = 3.0 * N * R
    axes = [Lk for Lk in L if Lk > 1]
    gemms = [2] + [4] * (len(axes) - 1)          # the first axis has a zero imaginary part
    flops = sum(g * 2.0 * nr3 * Lk for g, Lk in zip(gemms, axes))
    ms = dev_ms / max(n, 1)
    # bytes each stage must move once (replica pitch Rp = R here): the first
    # axis reads the 3 spin planes and writes U re/im, every later axis reads
    # U re/im and writes U re/im (the last one the norm plane instead), and
    # the transpose reads the norm plane and writes the [replica][q] result
    plane = 4.0 * N * R
    nbytes = 3 * plane + 6 * plane                       # first axis
    nbytes += (len(axes) - 2) * 12 * plane if len(axes) > 2 else 0.0
    nbytes += 6 * plane + plane if len(axes) > 1 else 0.0
    nbytes += 2 * plane                                  # transpose
    return {"workload": f"tri_afm_96: {L[0]}x{L[1]} cells ({N} spins, {N} q points), {R} replicas",
            "device_ms_per_call": ms, "wall_ms_per_call": wall * 1e3, "calls": n,
            "roofline": {"bound": "hbm", "traffic_bytes_per_call": nbytes,
                         "achieved": nbytes / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "note": "bytes every stage must read and write once over the device time of a call; "
                                 "the MFMA issue rate below is far from its peak, the stages move bytes"},
            "dft_flops": flops, "mfma_flops": 3.0 * flops,
            "mfma_tflops_per_s": 3.0 * flops / (ms * 1e-3) / 1e12,
            "direct_sum_flops": 12.0 * N * N * R,
            "direct_sum_equivalent_tflops_per_s": 12.0 * N * N * R / (ms * 1e-3) / 1e12,
            "max_error_vs_fp64_fft": worst,
            "error_note": "max |S_gpu - S_fp64| / max S_fp64 over 4 replicas"}


def batch_bench(mcs, n_lattices=256, kinds=None, plain=True):
    """The batched uneven workload: 256 lattices (simple-cubic Ising and
    Heisenberg, triangular, pyrochlore; 5.35 M spins, 8-32 replicas each,
    114 M spin updates per MCS), each with its own tempering ladder (swaps
    every 5 MCS) and a sample every 10 MCS.  Batched: two batches (Ising,
    Heisenberg) issue one launch per kernel kind and template for all their
    lattices, whole periods replayed as graphs.  Plain: the same engines run
    one after another with ether_run (one launch per lattice per kernel, each
    with its own graphs).  Both time `mcs` MCS after a 10-MCS warmup.
    `kinds` keeps only those lattice kinds and plain=False skips the plain
    leg (for profiles of the batched kernels)."""
    import numpy as np
    from ether_amd import engine as E
    from ether_amd import workloads
    from ether_amd.driver import attach_neighbours, make_engine
    from concurrent.futures import ThreadPoolExecutor
    t_setup = time.perf_counter()
    specs = [sp for sp in workloads.c5_lattices(1992, n_lattices) if not kinds or sp["kind"] in kinds]

    def one(sp):
        m = workloads.c5_model(sp)
        attach_neighbours(m)
        eng = make_engine(m, sp["R"], seed=sp["seed"])
        eng.set_betas(m.betas(workloads.c5_temperatures(sp)))
        eng.set_ladders([0, sp["R"]])
        eng.init_spins(E.INIT_RANDOM)
        return m, eng

    # contexts are independent: their set-up (neighbours, commit) runs in
    # native code on several host threads at once
    engines, updates = {"ising": [], "xyz": []}, 0.0
    with ThreadPoolExecutor(max_workers=8) as pool:
        for sp, (m, eng) in zip(specs, pool.map(one, specs)):
            engines["ising" if m.tags.ising else "xyz"].append(eng)
            updates += float(m.n_sites) * sp["R"]
    setup_s = time.perf_counter() - t_setup
    allg = engines["ising"] + engines["xyz"]
    batches = [E.Batch(v) for v in engines.values() if v]
    sched = dict(mcs_eq=0, coa=10, pt=5)

    swaps = [0] * len(batches)

    def run_batched(first, n):
        # each batch's swap count carried from call to call, as a driver does
        for k, b in enumerate(batches):
            swaps[k] = b.run(first, n, sched["mcs_eq"], sched["coa"], 0, 0.0, sched["pt"], swaps[k])

    swaps_plain = [0] * len(allg)

    def run_plain(first, n):
        for k, e in enumerate(allg):
            swaps_plain[k] = e.run(first, n, sched["mcs_eq"], sched["coa"], 0, 0.0, sched["pt"], swaps_plain[k])

    def sync():
        for e in allg:
            e.synchronize()

    step = 1
    run_batched(step, 10)
    step += 10
    sync()
    t0 = time.perf_counter()
    run_batched(step, mcs)
    sync()
    t_batched = time.perf_counter() - t0
    step += mcs
    t_plain = None
    if plain:
        run_plain(step, 10)
        step += 10
        sync()
        t0 = time.perf_counter()
        run_plain(step, mcs)
        sync()
        t_plain = time.perf_counter() - t0
    for b in batches:
        b.close()
    for e in allg:
        e.close()
    return {"workload": f"{len(specs)} lattices (simple-cubic Ising/Heisenberg 8..48, triangular "
                        f"24..128, pyrochlore 4..12 cells), {int(updates)} spin updates per MCS, own "
                        f"ladders of 8-32 replicas, swaps every 5 MCS, sample every 10 MCS; {mcs} MCS",
            "batched_spin_updates_per_s": updates * mcs / t_batched,
            "plain_spin_updates_per_s": updates * mcs / t_plain if t_plain else None,
            "speedup": t_plain / t_batched if t_plain else None,
            "batched_ms_per_mcs": t_batched / mcs * 1e3,
            "plain_ms_per_mcs": t_plain / mcs * 1e3 if t_plain else None,
            "lattices": {k: len(v) for k, v in engines.items()}, "setup_s": setup_s}


def pt_brief(cmp):
    """Round trips, round trips per second, round-trip cost and the lowest
    swap acceptance per ladder of one pt_comparison."""
    def one(r):
        return {"round_trips": r["round_trips"], "per_s": round(r["round_trips_per_s"], 4),
                "cost": round(r["round_trip_cost"], 2), "swap_acc_min": round(r["swap_acc_min"], 5),
                "spin_updates_per_s": float(f"{r['spin_updates_per_s']:.4g}")}
    return {"workload": cmp["workload"].split(";")[0],
            "linear": one(cmp["linear"]), "learned": one(cmp["learned"]),
            "plain_per_ladder": one(cmp["plain_per_ladder"]),
            "learned_ladder_kept": cmp["learned"]["ladder"]}


def pt_comparison(mcs, cells=16, temp=None):
    """Pyrochlore tempering: linear ladder versus the learned ladder.

    pyro_pt_16: cells^3 cubic cells x 16 sites (65536 spins at 16^3), J1 = 1,
    J2 = 0.05 meV, 64 temperatures (1 -> 0.05 K unless `temp` gives another
    TEMP tag), swaps every 11 MCS, REPEAT 4 (256 chains); the MLP is trained
    on <e> and the acceptance-consistent C measured in the first 2000 MCS
    (four windows, retuned after each).
    Reports completed cold->hot->cold round trips per second of the production
    loop and the spin updates per second of the same loop."""
    from ether_amd import ladder, workloads
    over = {"sc": f"{cells} {cells} {cells}"}
    if temp:
        over["temp"] = temp
    m = workloads.model("pyro_pt_16", **over)
    temps = m.tags.temperatures()
    t0 = time.perf_counter()
    lin, lrn = ladder.compare_ladders(m, train_mcs=2000, mcs=mcs)
    plain = ladder.plain_run(m, train_mcs=2000, mcs=mcs)
    total = time.perf_counter() - t0

    def rep(run):
        return {"round_trips": run.round_trips, "wall_s": run.wall_s, "mcs": run.mcs,
                "round_trips_per_s": run.round_trips_per_s,
                "spin_updates_per_s": run.updates_per_s,
                "swap_acc_min": float(run.acc_rate.min()), "swap_acc_max": float(run.acc_rate.max()),
                "swap_acc_cv": float(run.acc_rate.std() / max(run.acc_rate.mean(), 1e-30)),
                "ladder": run.source,
                "round_trip_cost": run.cost,
                "metropolis_acc_min": float(run.metro_acc.min()),
                "metropolis_acc_max": float(run.metro_acc.max()),
                "swap_acc": [round(float(x), 4) for x in run.acc_rate],
                "temps": [round(float(x), 6) for x in run.temps]}

    return {"workload": f"pyrochlore {cells}^3x16 ({m.n_sites} spins), J1=1 J2=0.05 meV, "
                        f"{len(temps)} T {temps[0]:g}->{temps[-1]:g} K x {m.tags.repeat} repeats, "
                        f"swaps every {m.tags.exchange_interval} MCS, MCS {mcs}; "
                        f"2000 training MCS, then the ladder under test",
            "linear": rep(lin), "learned": rep(lrn),
            "plain_per_ladder": rep(plain),
            "plain_note": "one context per repeat (64 replicas each), linear ladder, launched "
                          "one lattice at a time; linear/learned run all repeats in one context",
            "mlp_final_loss": lrn.mlp_loss,
            "gain_round_trips_per_s": (lrn.round_trips_per_s / lin.round_trips_per_s
                                       if lin.round_trips_per_s > 0 else None),
            "total_s": total}


if __name__ == "__main__":
    main()
