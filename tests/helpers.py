"""Shared set-up for the GPU parity tests: one engine and one oracle on the
same lattice, model and seed."""
import sys

import numpy as np

from ether_amd import workloads
from ether_amd.driver import attach_neighbours, make_engine
from oracle import oracle as o


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pair(name, n_replicas, seed=1992, **over):
    m = workloads.model(name, **over)
    attach_neighbours(m)
    eng = make_engine(m, n_replicas, seed=seed)
    orc = o.Oracle.from_model(m)
    return m, eng, orc


def colour_order(eng):
    """Sites sorted by colour (the device's phase order) and their RNG ids."""
    nc, col, rid = eng.layout()
    order = np.argsort(col * (eng.N + 1) + np.arange(eng.N), kind="stable").astype(np.int32)
    return nc, col, order, rid


def random_xyz(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def assert_runs_agree(mean_a, sem_a, mean_b, sem_b, n_a, n_b, label):
    """Whole-run estimators of two samplers (means over n_a and n_b repeats,
    SEMs from the repeats) agree: Welch's t per comparison; the smallest
    p-value above 0.001 / K (K comparisons: a family-wise 0.1% test) and no
    more p < 0.01 than 1% of K plus three binomial standard deviations (at
    least 2).  Estimators that are constant in both (a fully ordered Ising
    scan) get a floor of 1e-9 of the observable's scale; NaN (U where the
    order parameter vanishes) is skipped."""
    import math
    from scipy import stats
    scale = np.maximum(np.abs(mean_a).max(axis=0), np.abs(mean_b).max(axis=0))
    floor = 1e-9 * (np.abs(mean_a) + np.abs(mean_b) + scale[None, :]) + 1e-300
    va, vb = sem_a ** 2, sem_b ** 2
    se2 = va + vb + floor ** 2
    t = (mean_a - mean_b) / np.sqrt(se2)
    nu = se2 ** 2 / (va ** 2 / (n_a - 1) + vb ** 2 / (n_b - 1) + 1e-300)
    ok = np.isfinite(t)
    p = 2.0 * stats.t.sf(np.abs(t[ok]), np.maximum(nu[ok], 1.0))
    K = p.size
    low = int(np.sum(p < 0.01))
    allowed = max(2, math.ceil(0.01 * K + 3.0 * math.sqrt(0.01 * 0.99 * K)))
    say(f"{label}: {K} comparisons, min p {p.min():.2e} (bound {1e-3 / K:.1e}), "
        f"{low} below 0.01 (allowed {allowed}), max |t| {np.abs(t[ok]).max():.2f}")
    assert p.min() > 1e-3 / K
    assert low <= allowed


def three_shell_ising(L):
    """Simple-cubic Ising with three coupling classes: J = -1, 0.5, 0.25 meV
    on the 6, 12 and 8 neighbours at 1, sqrt 2, sqrt 3 (26 per site), with
    its neighbour lists attached."""
    from ether_amd.inputs import build_model, parse_in_ether, parse_j_exchange, parse_structure
    st = parse_structure(workloads.SC_STRUCTURE)
    tags = parse_in_ether(f"model = ising\nmcs = 100 10\ntemp = 2 2 1\nspin = 1\nspecies = A\n"
                          f"sc = {L} {L} {L}\nbc = c c c\nrepeat = 1\ncoa = 10\nseed = 1992\n"
                          f"para = .T. 1\n", st.n_species)
    ex = parse_j_exchange("3\n1.0 1\nA A -1 -1 -1\n1.414214 1\nA A 0.5 0.5 0.5\n"
                          "1.732051 1\nA A 0.25 0.25 0.25\n")
    m = build_model(tags, st, ex, None, None)
    attach_neighbours(m)
    return m
