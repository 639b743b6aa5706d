"""The structure factor on the GPU for the reference's inputs
(utility/Structure_Factor/structure_factor.f90:99-177), against the oracle's
FP64 direct sum.

Grid form with several sites per cell (ether_structure_factor_basis): the
pyrochlore cell (16 sites) and the reference's SimpleCubic FM example (two
species, 4 + 4 sites in a 2 A cell), every q of the commensurate grid of the
supercell, q = 2 pi sum_k m_k b_k / L_k.

Line form (ether_structure_factor_line): the utility's own q-line,
q_m = q_start + m dq along k = b1 + b2 + b3 (reciprocal vectors with 2 pi),
every q against the oracle's whole output line {sum_a |S_a|^2/N^2, |S_x|/N,
|S_y|/N, |S_z|/N}, including phases q k.r of several hundred radians (the
device forms them in FP64 before its FP32 sine and cosine), and Ising spins.

Tolerances: split-BF16 products keep about 16 mantissa bits and FP32
accumulation adds its rounding, so |S_a|/N is within 5e-5 absolute and the
sum of |S_a|^2/N^2 within 1e-4 of the largest value of the replica plus 1e-6.
"""
import math
import os

import numpy as np
import pytest

from ether_amd.driver import make_engine, attach_neighbours
from ether_amd.inputs import read_directory
from tests.helpers import pair, random_xyz, say
from oracle import oracle as o

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "tests", "golden", "examples")


def _states(m, eng, R, seed):
    """Random states, one with a Bragg peak (an up/down pattern by sublattice)
    and one ferromagnet."""
    rng = np.random.default_rng(seed)
    out = []
    nb = m.basis_xyz.shape[0]
    for r in range(R):
        if r % 3 == 0:
            S = random_xyz(rng, m.n_sites)
        elif r % 3 == 1:
            sign = np.where(np.arange(m.n_sites) % nb < nb // 2, 1.0, -1.0)
            S = np.outer(sign, [0.0, 0.6, 0.8]) + 0.1 * random_xyz(rng, m.n_sites)
            S /= np.linalg.norm(S, axis=1)[:, None]
        else:
            S = np.tile([0.8, 0.0, 0.6], (m.n_sites, 1))
        S = S.astype(np.float32).astype(np.float64)
        eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))
        out.append(S)
    return out


def _basis_case(m, R):
    attach_neighbours(m)
    eng = make_engine(m, R)
    eng.set_betas(m.betas([5.0] * R))
    states = _states(m, eng, R, 21)
    abc = m.structure.abc
    frac = m.basis_xyz @ np.linalg.inv(abc)
    nb = frac.shape[0]
    # the supercell order puts the basis innermost: site l of cell 0 is basis site l
    assert np.allclose(m.positions[:nb], m.basis_xyz)
    L = m.sc
    sq = eng.structure_factor_basis(L, frac)
    b = np.linalg.inv(abc).T
    worst = 0.0
    for r in range(R):
        ref = []
        for m3 in range(L[2]):
            for m2 in range(L[1]):
                for m1 in range(L[0]):
                    q = 2 * math.pi * (m1 * b[0] / L[0] + m2 * b[1] / L[1] + m3 * b[2] / L[2])
                    ref.append(o.structure_factor(m.positions, states[r], q))
        ref = np.array(ref)
        err = np.abs(sq[r] - ref)
        worst = max(worst, float(err.max() / max(ref.max(), 1e-30)))
        assert np.all(err <= 1e-4 * ref.max() + 1e-6), (r, err.max(), ref.max())
        if r % 3 == 2:
            assert sq[r][0] == pytest.approx(1.0, abs=1e-4)      # uniform state: S(0) = 1
    eng.close()
    return worst


def test_structure_factor_basis_pyrochlore():
    say("sq basis: pyrochlore 3x2x2 cells x 16 sites")
    from ether_amd import workloads
    m = workloads.model("pyro_pt_16", sc="3 2 2", pt=".F. 11")
    worst = _basis_case(m, 6)
    say(f"pyrochlore grid: worst error {worst:.2e} of the peak")


def test_structure_factor_basis_reference_fm_example():
    say("sq basis: examples/SimpleCubic/FM, 4x4x4 cells x 8 sites")
    m = read_directory(os.path.join(EX, "SimpleCubic_FM"))
    assert m.basis_xyz.shape[0] == 8
    worst = _basis_case(m, 3)
    say(f"SimpleCubic FM grid: worst error {worst:.2e} of the peak")


def _kvec(abc):
    """k = b1 + b2 + b3 with b the reciprocal vectors, 2 pi included
    (structure_factor.f90:99-116)."""
    a = np.asarray(abc, np.float64)
    vol = np.dot(a[0], np.cross(a[1], a[2]))
    b = np.array([np.cross(a[1], a[2]), np.cross(a[2], a[0]), np.cross(a[0], a[1])]) * 2 * math.pi / vol
    return b.sum(axis=0)


@pytest.mark.parametrize("name,over,q,R", [
    ("pyro_pt_16", dict(sc="2 2 2", pt=".F. 11"), (0.0, 1.0, 0.05), 70),
    ("pyro_pt_16", dict(sc="4 4 4", pt=".F. 11"), (20.0, -5.0, 0.25), 3),     # large phases, falling q
    ("tri_afm_96", dict(sc="24 24 1"), (0.0, 2.0, 0.02), 5),
    ("sc_ising_16", dict(sc="8 8 8"), (0.0, 1.0, 0.125), 4),
])
def test_structure_factor_line_matches_direct_sum(name, over, q, R):
    say(f"sq line {name} {over} q={q} R={R}")
    m, eng, _ = pair(name, R, **over)
    eng.set_betas(m.betas([5.0] * R))
    if m.tags.ising:
        rng = np.random.default_rng(5)
        states = []
        for r in range(R):
            S = np.zeros((m.n_sites, 3))
            S[:, 2] = np.where(rng.random(m.n_sites) < (0.5 if r % 2 else 0.9), 1.0, -1.0)
            eng.set_spins(r, S)
            states.append(S)
    else:
        states = _states(m, eng, R, 8)
    k = _kvec(m.structure.abc)
    q_start, q_stop, dq = q
    out = eng.structure_factor_line(m.positions, k, q_start, q_stop, dq)
    n_q = abs(int(round((q_stop - q_start) / dq))) + 1
    assert out.shape == (R, n_q, 4)
    phase_max = 0.0
    for r in range(R):
        ref = np.array([o.structure_factor_parts(m.positions, states[r], (q_start + j * dq) * k)
                        for j in range(n_q)])
        phase_max = max(phase_max, float(np.abs(np.outer(q_start + dq * np.arange(n_q), m.positions @ k)).max()))
        assert np.all(np.abs(out[r, :, 1:] - ref[:, 1:]) <= 5e-5), (r, np.abs(out[r, :, 1:] - ref[:, 1:]).max())
        assert np.all(np.abs(out[r, :, 0] - ref[:, 0]) <= 1e-4 * ref[:, 0].max() + 1e-6), r
        if m.tags.ising:
            assert not out[r, :, 1:3].any()
    say(f"largest phase {phase_max:.0f} rad")
    eng.close()


@pytest.mark.parametrize("name,over,R,basis", [
    ("tri_afm_96", {}, 70, False),                                 # 96 x 96: the benchmark's axis length
    ("sc_heis_afm_32", dict(sc="12 6 1"), 130, False),
    ("sc_heis_afm_32", dict(sc="4 4 4"), 3, False),                # three axes of one block of rows
    ("sc_heis_afm_32", dict(sc="40 1 1"), 5, False),               # one axis: first and last stage at once
    ("sc_heis_afm_32", dict(sc="100 3 1"), 3, False),              # an axis above 96 keeps the tiled kernel
    ("sc_ising_16", dict(sc="8 8 8"), 9, False),                   # the int8 Ising plane
    ("pyro_pt_16", dict(sc="3 2 2", pt=".F. 11"), 6, True),        # 16 sites per cell, several W sets
])
def test_structure_factor_axis_kernels_bit_identical(name, over, R, basis):
    """The grid form with the axis's DFT matrices resident in LDS (one wave
    per line and 16 replicas) against the tiled kernel that stages both
    operands per K step (ether_set_dft_resident(ctx, 0)): the same split-BF16
    products in the same order per element, so S(q) must be bit-identical
    (the tiled kernel is itself checked against the FP64 direct sum and FFT
    above)."""
    say(f"sq axis kernels {name} {over} R={R}")
    m, eng, _ = pair(name, R, **over)
    eng.set_betas(m.betas([5.0] * R))
    rng = np.random.default_rng(31)
    for r in range(R):
        if m.tags.ising:
            S = np.zeros((m.n_sites, 3))
            S[:, 2] = np.where(rng.random(m.n_sites) < 0.3 + 0.4 * (r % 2), 1.0, -1.0)
        else:
            S = random_xyz(rng, m.n_sites).astype(np.float32).astype(np.float64)
        eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))
    if basis:
        frac = m.basis_xyz @ np.linalg.inv(m.structure.abc)
        run = lambda: eng.structure_factor_basis(m.sc, frac)   # noqa: E731
    else:
        run = lambda: eng.structure_factor(m.sc)               # noqa: E731
    resident = run()
    eng.set_dft_resident(False)
    tiled = run()
    eng.set_dft_resident(True)
    again = run()
    assert resident.shape == tiled.shape
    assert np.array_equal(resident, tiled), float(np.abs(resident - tiled).max())
    assert np.array_equal(resident, again)
    eng.close()
