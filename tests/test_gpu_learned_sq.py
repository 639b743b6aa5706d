"""GPU checks of the learned-ladder MLP kernels and the MFMA structure factor.

MLP: the device forward pass and the gradient of the weighted squared loss
(hand-written backward kernel) against the FP64 oracle, whose own gradient is
pinned by finite differences; training lowers the loss; the ladder placement
equalises the thermodynamic length between neighbours.

S(q): the split-BF16 MFMA contraction against the oracle's direct FP64 sum
exp(i q.r) (utility/Structure_Factor/structure_factor.f90:150-177) at every
commensurate q of small lattices; the error bound is stated in the test.
"""
import math

import numpy as np
import pytest

from tests.helpers import pair, random_xyz, say
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def _mlp_engine():
    m, eng, _ = pair("sc_ising_16", 1, sc="2 2 2")
    return eng


def test_mlp_forward_and_gradient_match_oracle():
    say("mlp grad")
    eng = _mlp_engine()
    eng.mlp_init(5)
    p = eng.mlp_params().astype(np.float64)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, 48).astype(np.float32)
    y = rng.normal(size=(48, 2)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, (48, 2)).astype(np.float32)
    yf = eng.mlp_forward(x)
    yo = o.mlp_forward(p, x.astype(np.float64))
    assert np.max(np.abs(yf - yo)) < 2e-5 * (1 + np.abs(yo).max())
    loss, g = eng.mlp_loss_grad(x, y, w)
    lo, go = o.mlp_loss_grad(p, x.astype(np.float64), y.astype(np.float64), w.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-4)
    assert np.max(np.abs(g - go)) <= 1e-4 * np.abs(go).max()
    eng.close()


def test_mlp_training_fits_a_heat_capacity_curve():
    say("mlp train")
    eng = _mlp_engine()
    eng.mlp_init(1)
    T = np.linspace(0.5, 3.0, 64)
    cv = 1.0 / (1.0 + ((T - 1.5) / 0.2) ** 2) + 0.3      # a peaked C_v
    e = -np.tanh(2.0 / T)
    x = ((T - 1.75) / 1.25).astype(np.float32)
    ly = np.log(cv)
    y = np.stack([(e - e.mean()) / e.std(), (ly - ly.mean()) / ly.std()], axis=1).astype(np.float32)
    w = np.ones_like(y)
    l0, _ = eng.mlp_loss_grad(x, y, w)
    loss = eng.mlp_train(x, y, w, 3000, 3e-3)
    assert loss < 0.02 * l0
    pred = eng.mlp_forward(x)
    assert np.max(np.abs(pred - y)) < 0.15
    # ladder: equal int sqrt(C_v)/T dT between neighbours, hottest first
    Tl = eng.ladder_place(1.75, 1.25, ly.std(), 0.5, 3.0, 16)
    assert Tl[0] == pytest.approx(3.0) and Tl[-1] == pytest.approx(0.5)
    assert np.all(np.diff(Tl) < 0)
    grid = np.linspace(0.5, 3.0, 20001)
    f = np.sqrt(np.interp(grid, T, cv)) / grid
    cum = np.concatenate([[0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(grid))])
    seg = np.diff(np.interp(Tl[::-1], grid, cum))
    assert seg.std() / seg.mean() < 0.08
    eng.close()


@pytest.mark.parametrize("L", [(8, 8, 1), (12, 6, 1), (4, 4, 4)])
def test_structure_factor_matches_direct_sum(L):
    say(f"sq {L}")
    sc = f"{L[0]} {L[1]} {L[2]}"
    R = 3
    m, eng, _ = pair("sc_heis_afm_32", R, sc=sc)
    eng.set_betas(m.betas([5.0] * R))
    rng = np.random.default_rng(4)
    states = []
    for r in range(R):
        S = random_xyz(rng, m.n_sites)
        if r == 1:   # a Neel-like state with a Bragg peak
            sign = np.where(((np.arange(m.n_sites) % L[0]) + (np.arange(m.n_sites) // L[0])) % 2 == 0, 1, -1)
            S = 0.9 * np.outer(sign, [0, 0, 1.0]) + 0.1 * S
            S /= np.linalg.norm(S, axis=1)[:, None]
        S = S.astype(np.float32).astype(np.float64)
        eng.set_spins(r, S)
        states.append(S)
    sq = eng.structure_factor(L)
    abc = m.structure.abc
    b = np.linalg.inv(abc).T          # rows b_k with a_i . b_j = delta_ij
    N = m.n_sites
    worst = 0.0
    for r in range(R):
        S = states[r]
        ref = np.zeros(N)
        q_idx = 0
        for m3 in range(L[2]):
            for m2 in range(L[1]):
                for m1 in range(L[0]):
                    q = 2 * math.pi * (m1 * b[0] / L[0] + m2 * b[1] / L[1] + m3 * b[2] / L[2])
                    ref[q_idx] = o.structure_factor(m.positions, S, q)
                    q_idx += 1
        err = np.abs(sq[r] - ref)
        worst = max(worst, err.max())
        # split-BF16 with FP32 accumulation: ~2^-16 per product, summed over N terms;
        # bound |dS|^2 error by 1e-4 of the largest peak plus 1e-6 absolute
        assert np.all(err <= 1e-4 * max(ref.max(), 1.0 / N) + 1e-6), (r, err.max(), ref.max())
    say(f"S(q) worst abs error vs FP64: {worst:.3e}")
    eng.close()


def test_structure_factor_triangular_96_against_fp64_fft():
    """S(q) at the triangular benchmark size (96 x 96 cells, one site each,
    9216 q points on the commensurate grid) for 64 replicas, against numpy's
    FP64 FFT: on a one-site-per-cell supercell q.r = 2 pi (m1 n1/L1 + m2 n2/L2),
    so S_a(q) is the 2-D DFT of spin component a over the cell indices, and
    |S_a(q)|^2 does not depend on the sign convention for real spins.  States:
    random, a 120-degree order with a Bragg peak at K, and a ferromagnet with
    S(0) = 1.  Split-BF16 products (hi*hi + hi*lo + lo*hi) accumulated in FP32
    drop lo*lo, about 2^-16 of each product; on a uniform state those errors add
    up coherently, so the bound is 1e-4 of the largest S(q) of the replica.  The
    worst error seen is printed."""
    say("sq triangular 96x96 vs FP64 FFT")
    R = 64
    m, eng, _ = pair("tri_afm_96", R)
    assert m.n_sites == 96 * 96
    eng.set_betas(m.betas([5.0] * R))
    rng = np.random.default_rng(11)
    L1, L2 = 96, 96
    n1 = np.arange(m.n_sites) % L1
    n2 = np.arange(m.n_sites) // L1
    states = []
    for r in range(R):
        if r % 3 == 0:
            S = random_xyz(rng, m.n_sites)
        elif r % 3 == 1:
            ang = 2 * np.pi / 3 * ((n1 + 2 * n2) % 3)          # three-sublattice 120-degree state
            S = np.stack([np.cos(ang), np.sin(ang), 0.05 * rng.normal(size=m.n_sites)], axis=1)
            S /= np.linalg.norm(S, axis=1)[:, None]
        else:
            S = np.tile([0.0, 0.6, 0.8], (m.n_sites, 1))
        S = S.astype(np.float32).astype(np.float64)
        eng.set_spins(r, S, np.arctan2(S[:, 1], S[:, 0]))
        states.append(S)
    sq = eng.structure_factor((L1, L2, 1))
    worst_rel = 0.0
    for r in range(R):
        grid = states[r].reshape(L2, L1, 3)                     # [n2][n1][a]
        ref = np.zeros((L2, L1))
        for a in range(3):
            ref += np.abs(np.fft.fft2(grid[:, :, a])) ** 2
        ref = ref.reshape(-1) / m.n_sites ** 2                  # q index m1 + L1*m2
        err = np.abs(sq[r].astype(np.float64) - ref).max()
        rel = err / ref.max()
        worst_rel = max(worst_rel, rel)
        assert rel <= 1e-4, (r, err, ref.max())
        if r % 3 == 2:
            assert abs(sq[r][0] - 1.0) < 1e-4                    # ferromagnet: S(0) = 1
    say(f"S(q) 96x96: worst error vs FP64 FFT {worst_rel:.2e} of the peak")
    eng.close()
