"""Distributional checks of the device sweeps against exact results.

Hundreds of independent replicas run on the device; the spread across
replicas gives the error bar, and every comparison is at 4 sigma:
  - 4x4 periodic Ising (2^16 states enumerated exactly): <e>, <e^2>, <|m|>;
  - one classical spin in a field: <S_z> = coth(x) - 1/x, x = beta g muB s H;
  - one spin with easy-axis anisotropy: <S_z^2> from the 1-D integral;
  - the same single spin with ANGLE moves.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from ether_amd import engine as E
from tests.helpers import pair, say

pytestmark = pytest.mark.gpu


def _exact_4x4(beta, J=-1.0):
    L, n = 4, 16
    es, ms = [], []
    for state in range(1 << n):
        s = np.array([1 if (state >> k) & 1 else -1 for k in range(n)]).reshape(L, L)
        es.append(J * (s * np.roll(s, 1, 0) + s * np.roll(s, 1, 1)).sum() / n)
        ms.append(abs(s.sum()) / n)
    es, ms = np.array(es), np.array(ms)
    w = np.exp(-beta * n * (es - es.min()))
    Z = w.sum()
    return (w * es).sum() / Z, (w * es * es).sum() / Z, (w * ms).sum() / Z


def _per_replica(mom):
    n = mom[:, E.MOM_N]
    return mom[:, E.MOM_E] / n, mom[:, E.MOM_E2] / n, mom[:, E.MOM_M] / n


def _check(x, exact, what):
    mean, se = x.mean(), x.std(ddof=1) / math.sqrt(len(x))
    assert abs(mean - exact) < 4 * se + 1e-12, f"{what}: {mean} vs {exact} (se {se})"


def test_ising_4x4_exact_enumeration():
    say("stats ising 4x4")
    R = 256
    m, eng, _ = pair("sc_ising_16", R, sc="4 4 1", para=".T. 1")
    beta = 0.4
    eng.set_betas([beta] * R)
    eng.init_spins(E.INIT_RANDOM)
    eng.reset_moments()
    eng.run(1, 2200, 200, 1)
    e, e2, mm = _per_replica(eng.moments())
    ee, ee2, em = _exact_4x4(beta)
    _check(e, ee, "e")
    _check(e2, ee2, "e^2")
    _check(mm, em, "|m|")
    eng.close()


@pytest.mark.parametrize("angle", [False, True])
def test_single_spin_in_field_langevin(angle):
    say(f"stats langevin angle={angle}")
    R = 4096
    m, eng, _ = pair("sc_heis_afm_32", R, sc="1 1 1", para=".T. 1", zeeman=".T. 0 0 1.7",
                     cb=".F.", ovrr=".F. 0 0", angle=".T. 40" if angle else ".F. 0")
    assert m.n_sites == 1 and m.nn_count.sum() == 0
    beta = 1.0
    eng.set_betas([beta] * R)
    eng.init_spins(E.INIT_RANDOM)
    eng.reset_moments()
    eng.run(1, 600, 100, 1)
    mom = eng.moments()
    # m = |M|/N is |S|=1; use the energy: e = E_tot/(2N) = -g muB s H S_z / 2
    ez = mom[:, E.MOM_E] / mom[:, E.MOM_N]
    sz = -2.0 * ez / 1.7
    x = beta * 1.7
    _check(sz, 1 / math.tanh(x) - 1 / x, "S_z")
    eng.close()


def test_single_spin_easy_axis():
    say("stats sia")
    R = 4096
    m, eng, _ = pair("tri_afm_96", R, sc="1 1 1", para=".T. 1", zeeman=".F. 0 0 0",
                     ovrr=".F. 0 0", angle=".F. 0")
    assert m.n_sites == 1
    D = m.sia_vec[0, 2]          # -0.1 in PARA units
    beta = 20.0
    eng.set_betas([beta] * R)
    eng.init_spins(E.INIT_RANDOM)
    eng.reset_moments()
    eng.run(1, 600, 100, 1)
    mom = eng.moments()
    e = mom[:, E.MOM_E] / mom[:, E.MOM_N]
    z2 = 2.0 * e / D             # e = D S_z^2 / 2
    num = quad(lambda t: t * t * math.exp(-beta * D * t * t), -1, 1)[0]
    den = quad(lambda t: math.exp(-beta * D * t * t), -1, 1)[0]
    _check(z2, num / den, "S_z^2")
    eng.close()
