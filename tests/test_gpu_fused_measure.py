"""The measurement booked by the sweep (XYZ).

When a measurement follows a sweep (a sampling step, or a swap step), the
sweep's colour phases add each site's terms as they update it: the exchange
energy of its bonds to neighbours of earlier colours (final already), each
bond booked once from its later-coloured end and counted twice, and its
Zeeman, SIA and magnetisation terms.  With two colours a phase's neighbours
are all earlier or all later and the workgroups add their sums to shared
fixed-point totals; with three or more (triangular, pyrochlore) a phase
books the field of its earlier neighbours alone and every workgroup writes a
record of its own, which the finalize adds up.  Lattices whose bonds lack
their mirror measure in a pass.  Checked here against the standalone measurement pass and the
oracle's FP64 sums (src/get_tot_energy.f90:35-45,
src/get_tot_magnetisation.f90:35-38) on the state the sweep leaves, to FP32
accuracy (1e-6 of the sum of |site energies|), and for independence of the
launch shape: every term is rounded per site to a fixed-point integer and
summed exactly, so sites per wave, the XCD map and lanes per site give
bit-identical moments.
"""
import numpy as np
import pytest

from ether_amd import engine as E
from tests.helpers import pair, random_xyz, say

pytestmark = pytest.mark.gpu


def _start(m, eng, R, seed):
    rng = np.random.default_rng(seed)
    for r in range(R):
        S = random_xyz(rng, m.n_sites)
        eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))


def _fields_pair(R, **over):
    """A 6 x 6 x 6 simple-cubic antiferromagnet (two colours) with a tilted
    Zeeman field, SIA and ANGLE moves: every word of the record is nonzero."""
    from ether_amd import workloads
    from ether_amd.driver import attach_neighbours, make_engine
    from ether_amd.inputs import build_model, parse_in_ether, parse_j_exchange, parse_sia, parse_structure
    from oracle import oracle as o
    tags = dict(sc="6 6 6", zeeman=".T. 0.3 0 1", g_factor="2", sia=".T.", angle=".T. 5")
    tags.update(over)
    f = workloads.files("sc_heis_afm_32", **tags)
    st = parse_structure(f["structure"])
    t = parse_in_ether(f["in.ether"], st.n_species)
    m = build_model(t, st, parse_j_exchange(f["j_exchange"]), parse_sia("A 0.02 0 -0.1\n"), None)
    attach_neighbours(m)
    return m, make_engine(m, R, seed=5), o.Oracle.from_model(m)


def _case(name, over, R):
    return _fields_pair(R, **over) if name == "sc_fields" else pair(name, R, **over)


# (workload, overrides, replicas, lanes per site, OVRR, booked by the sweep)
CASES = [
    ("sc_heis_afm_32", dict(sc="8 8 8"), 70, 0, 1, True),      # 2 colours: no / all earlier neighbours
    ("sc_heis_afm_32", dict(sc="2 2 2"), 5, 0, 1, True),       # +x and -x neighbours are one site
    ("sc_heis_afm_32", dict(sc="8 8 8"), 13, 16, 1, True),     # lanes kernel, 4 sites per wave
    ("sc_fields", dict(), 9, 0, 0, True),                      # Zeeman, SIA, ANGLE
    ("sc_fields", dict(angle=".F. 0"), 21, 32, 0, True),
    ("tri_afm_96", dict(sc="12 12 1"), 13, 16, 0, True),       # 3 colours: per-workgroup records
    ("tri_afm_96", dict(sc="12 12 1"), 20, 20, 0, True),       # the same, 20 lanes: 3 sites per step
    ("sc_heis_afm_32", dict(sc="8 8 8"), 19, 19, 1, True),     # 19 lanes, 3 sites per step (atomics)
    ("sc_fields", dict(), 9, 9, 0, True),                      # 9 lanes, 7 sites per step
    ("tri_afm_96", dict(sc="12 12 1"), 70, 0, 0, True),        # two waves of replicas per site
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 24, 0, 1, True),   # 4 colours, 18 neighbours
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 24, 32, 1, True),  # the same, lanes kernel
]


@pytest.mark.parametrize("name,over,R,lanes,ovrr,booked", CASES)
def test_fused_measurement_matches_pass_and_oracle(name, over, R, lanes, ovrr, booked):
    say(f"fused measurement {name} {over} R={R} lanes={lanes}")
    m, eng, orc = _case(name, over, R)
    if lanes:
        eng.set_lanes(lanes)
    eng.set_betas(m.betas(np.linspace(5.0, 50.0, R)))
    _start(m, eng, R, 11)
    # steps 1..3 with coa 3: two sweeps (with OVRR where the model allows it),
    # then a sampling sweep whose phases book the measurement (two colours)
    eng.set_timing(True)
    eng.run(1, 3, 0, 3, 1 if ovrr else 0, 1.0 if ovrr else 0.0)
    nc = eng.layout()[0]
    assert eng.timing(E.TIME_SWEEP_MEAS)[1] == (nc if booked else 0)
    assert eng.timing(E.TIME_MEASURE)[1] == 1
    eng.set_timing(False)
    parts_f, _ = eng.energy_parts()
    mom_f = eng.moments().copy()
    slot = eng.pt_state()[0]
    eng.reset_moments()
    eng.measure(sample=True)                     # the standalone pass on the same state
    parts_s, _ = eng.energy_parts()
    mom_s = eng.moments()
    N = m.n_sites
    for r in range(R):
        S, _ = eng.get_spins(r)
        ex, ze, si = orc.total_energy(S)
        scale = sum(abs(orc.site_energy(S, i)) for i in range(N))
        tol = 1e-6 * scale + 1e-300
        for k, ref in enumerate((ex, ze, si)):
            assert abs(parts_f[r, k] - ref) <= tol, (r, k, parts_f[r, k], ref)
            assert abs(parts_f[r, k] - parts_s[r, k]) <= tol
        assert parts_f[r, 3] == pytest.approx(0.5 * parts_f[r, 0] + parts_f[r, 1] + parts_f[r, 2])
        Mo = orc.magnetisation(S)
        mf = mom_f[slot[r], E.MOM_M]
        assert mom_f[slot[r], E.MOM_N] == 1.0
        assert abs(mf - np.linalg.norm(Mo) / N) <= 1e-6
        assert abs(mf - mom_s[slot[r], E.MOM_M]) <= 1e-6
        e_f = mom_f[slot[r], E.MOM_E]
        assert abs(e_f - (ex + ze + si) / (2 * N)) <= tol / (2 * N)
    eng.close()


@pytest.mark.parametrize("name,over,R,lanes_list,ovrr", [
    ("sc_heis_afm_32", dict(sc="8 8 8"), 70, [0], 1),
    ("sc_fields", dict(), 13, [64, 16, 32, 0, 20, 13], 0),      # 20 lanes: 3 sites per step
    ("tri_afm_96", dict(sc="12 12 1"), 13, [64, 16, 32, 0, 13], 0),
    ("tri_afm_96", dict(sc="12 12 1"), 130, [0], 0),
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 24, [64, 32, 0, 25], 1),
])
def test_fused_measurement_does_not_depend_on_launch_shape(name, over, R, lanes_list, ovrr):
    say(f"fused measurement shapes {name} R={R}")
    ref = None
    for spw, xcd, ln in [(4, 1, lanes_list[0]), (1, 0, lanes_list[-1]), (16, 1, lanes_list[0]),
                         (2, 2, lanes_list[0]), (8, 2, lanes_list[-1])] + \
            [(8, 0, ln) for ln in lanes_list]:
        m, eng, _ = _case(name, over, R)
        eng.set_lanes(ln)
        eng.set_betas(m.betas(np.linspace(5.0, 50.0, R)))
        _start(m, eng, R, 5)
        eng.set_launch(spw, xcd)
        eng.run(1, 8, 0, 2, 3 if ovrr else 0, 1.0 if ovrr else 0.0)
        got = (eng.moments().copy(), eng.energy_parts()[0])
        eng.close()
        if ref is None:
            ref = got
        else:
            assert np.array_equal(got[0], ref[0]), f"moments differ: spw={spw} xcd={xcd} lanes={ln}"
            assert np.array_equal(got[1], ref[1])


def _edited(edit, R):
    """An 8 x 8 x 1 square lattice (sc, open along z) whose neighbour lists
    are edited before the engine and the oracle are built from them."""
    from ether_amd import workloads
    from ether_amd.driver import attach_neighbours, make_engine
    from oracle import oracle as o
    m = workloads.model("sc_heis_afm_32", sc="8 8 1")
    attach_neighbours(m)
    edit(m)
    return m, make_engine(m, R, seed=7), o.Oracle.from_model(m)


def _self_bond(m):
    # site i also lists itself
    for i in range(m.n_sites):
        m.nn_idx[0, i, m.nn_count[0, i]] = i
        m.nn_count[0, i] += 1


def _one_way(m):
    # site 0 drops its first neighbour j, whose own list still holds site 0
    m.nn_idx[0, 0, :-1] = m.nn_idx[0, 0, 1:].copy()
    m.nn_count[0, 0] -= 1


@pytest.mark.parametrize("edit", [_self_bond, _one_way])
def test_unmirrored_bonds_measure_apart(edit):
    """Booking a bond once and counting it twice needs every bond's mirror
    with the same coupling and no bond of a site to itself.  Lists that break
    this (possible through the C-ABI) leave the measurement to its own pass
    over the full rows, counted once, which is the reference's double sum
    whatever the lists hold."""
    say(f"unmirrored bonds {edit.__name__}")
    m, eng, orc = _edited(edit, 3)
    eng.set_betas(m.betas([5.0, 10.0, 20.0]))
    _start(m, eng, 3, 2)
    eng.set_timing(True)
    eng.run(1, 2, 0, 2, 1, 1.0)
    assert eng.timing(E.TIME_SWEEP_MEAS)[1] == 0
    assert eng.timing(E.TIME_MEASURE)[1] == 1
    eng.set_timing(False)
    parts, _ = eng.energy_parts()
    for r in range(3):
        S, _ = eng.get_spins(r)
        ex, _, _ = orc.total_energy(S)
        scale = sum(abs(orc.site_energy(S, i)) for i in range(m.n_sites))
        assert abs(parts[r, 0] - ex) <= 1e-6 * scale
    eng.close()


def test_swap_step_books_the_energies_it_swaps_on():
    """A swap step's sweep books the measurement the swap reads (E per
    replica): the energies equal the pass's on the state after the step, and
    the swap statistics equal those of the same run with the pass."""
    say("fused measurement on swap steps")
    runs = []
    for compat in (0, E.COMPAT_MEASURE_PASS):
        m, eng, orc = pair("sc_heis_afm_32", 24, sc="6 6 6")
        eng.set_compat(compat)
        eng.set_betas(m.betas(np.geomspace(30.0, 3.0, 24)))
        eng.set_ladders([0, 24])
        _start(m, eng, 24, 9)
        eng.set_timing(True)
        eng.run(1, 4, 100, 1000, 1, 1.0, 2)            # swaps at steps 2 and 4, no samples
        booked = eng.timing(E.TIME_SWEEP_MEAS)[1]
        eng.set_timing(False)
        parts, _ = eng.energy_parts()
        slot, acc, att, _ = eng.pt_state()
        E_s, _ = eng.measure()
        parts_s, _ = eng.energy_parts()
        for r in range(24):
            S, _ = eng.get_spins(r)
            scale = sum(abs(orc.site_energy(S, i)) for i in range(m.n_sites))
            assert abs(parts[r, 0] - parts_s[r, 0]) <= 1e-6 * scale
        runs.append((booked, slot.copy(), acc.copy(), att.copy()))
        eng.close()
    (b0, s0, a0, t0), (b1, s1, a1, t1) = runs
    assert b0 == 2 * 2 and b1 == 0                      # two swap steps x two colours
    assert np.array_equal(t0, t1) and t0.sum() > 0
    # the booked and the passed energies differ by FP32 rounding only, which
    # decides no swap here
    assert np.array_equal(s0, s1) and np.array_equal(a0, a1)


def test_booked_measurement_at_full_size():
    """The 32 x 32 x 32 antiferromagnet with 190 replicas (three waves of
    replicas per site, OVRR fused): the booked energies equal the pass's for
    every replica, and the oracle's FP64 sums on three of them.  With unit
    spins and one coupling, |e_i| <= sum_j |J_ij|, the site energy of the
    aligned state, so that state's sum of |site energies| bounds every
    state's."""
    say("fused measurement sc_heis_afm_32 32^3 R=190")
    R = 190
    m, eng, orc = pair("sc_heis_afm_32", R)
    N = m.n_sites
    assert N == 32 ** 3
    eng.set_betas(m.betas(np.linspace(5.0, 50.0, R)))
    _start(m, eng, R, 17)
    eng.set_timing(True)
    eng.run(1, 3, 0, 3, 1, 1.0)
    nc = eng.layout()[0]
    assert nc == 2 and eng.timing(E.TIME_SWEEP_MEAS)[1] == nc
    eng.set_timing(False)
    parts_f, _ = eng.energy_parts()
    mom_f = eng.moments().copy()
    slot = eng.pt_state()[0]
    eng.reset_moments()
    eng.measure(sample=True)
    parts_s, _ = eng.energy_parts()
    mom_s = eng.moments()
    up = np.zeros((N, 3))
    up[:, 2] = 1.0
    bound = sum(abs(orc.site_energy(up, i)) for i in range(N))
    assert bound > 0
    assert np.all(np.abs(parts_f[:, 0] - parts_s[:, 0]) <= 1e-6 * bound)
    assert np.all(np.abs(mom_f[slot, E.MOM_M] - mom_s[slot, E.MOM_M]) <= 1e-6)
    assert np.all(mom_f[slot, E.MOM_N] == 1.0)
    for r in (0, 95, R - 1):
        S, _ = eng.get_spins(r)
        ex, _, _ = orc.total_energy(S)
        scale = sum(abs(orc.site_energy(S, i)) for i in range(N))
        assert scale <= bound * (1 + 1e-12)
        assert abs(parts_f[r, 0] - ex) <= 1e-6 * scale, (r, parts_f[r, 0], ex)
        assert abs(mom_f[slot[r], E.MOM_M] - np.linalg.norm(orc.magnetisation(S)) / N) <= 1e-6
    eng.close()


def test_booked_measurement_at_the_pyrochlore_workload():
    """The 16^3-cell pyrochlore with 256 replicas (four colours, 18
    neighbours, per-workgroup records): the booked energies and moments equal
    the pass's for every replica and the oracle's FP64 sums on two of them.
    The aligned state bounds the sum of |site energies| as in the simple-cubic
    case (unit spins, site energy at most sum_j |J_ij|)."""
    say("fused measurement pyro_pt_16 16^3 x 16 R=256")
    R = 256
    m, eng, orc = pair("pyro_pt_16", R, pt=".F. 11")
    N = m.n_sites
    assert N == 65536
    eng.set_betas(m.betas(np.geomspace(1.0, 0.05, R)))
    _start(m, eng, R, 23)
    eng.set_timing(True)
    eng.run(1, 3, 0, 3)
    nc = eng.layout()[0]
    assert nc == 4 and eng.timing(E.TIME_SWEEP_MEAS)[1] == nc
    eng.set_timing(False)
    parts_f, _ = eng.energy_parts()
    mom_f = eng.moments().copy()
    slot = eng.pt_state()[0]
    eng.reset_moments()
    eng.measure(sample=True)
    parts_s, _ = eng.energy_parts()
    mom_s = eng.moments()
    up = np.zeros((N, 3))
    up[:, 2] = 1.0
    bound = sum(abs(orc.site_energy(up, i)) for i in range(N))
    assert bound > 0
    assert np.all(np.abs(parts_f[:, 0] - parts_s[:, 0]) <= 1e-6 * bound)
    assert np.all(np.abs(mom_f[slot, E.MOM_M] - mom_s[slot, E.MOM_M]) <= 1e-6)
    assert np.all(mom_f[slot, E.MOM_N] == 1.0)
    for r in (0, R - 1):
        S, _ = eng.get_spins(r)
        ex, _, _ = orc.total_energy(S)
        assert abs(parts_f[r, 0] - ex) <= 1e-6 * bound, (r, parts_f[r, 0], ex)
        assert abs(mom_f[slot[r], E.MOM_M] - np.linalg.norm(orc.magnetisation(S)) / N) <= 1e-6
    eng.close()


@pytest.mark.parametrize("name,over,R", [
    ("tri_afm_96", dict(sc="12 12 1"), 70),
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), 24),
])
def test_records_survive_shape_changes(name, over, R):
    """Per-workgroup records (three or more colours): a run whose launch
    shape changes between measuring sweeps (fewer workgroups after more)
    gives the moments of a run at one shape, so no record of an earlier
    shape is read again."""
    say(f"measurement records across shapes {name} R={R}")
    got = []
    for shapes in ([(4, 1)] * 3, [(1, 0), (16, 1), (8, 2)]):
        m, eng, _ = _case(name, over, R)
        eng.set_betas(m.betas(np.linspace(5.0, 50.0, R)))
        _start(m, eng, R, 3)
        step = 1
        for spw, xcd in shapes:
            eng.set_launch(spw, xcd)
            eng.run(step, 4, 0, 2)
            step += 4
        got.append((eng.moments().copy(), eng.energy_parts()[0]))
        eng.close()
    assert np.array_equal(got[0][0], got[1][0])
    assert np.array_equal(got[0][1], got[1][1])
