"""The neighbour lists and the periodic colouring built on the GPU
(ether_build_neighbours_device) against the host builder, which the CPU tests
pin to the oracle's restatement of the reference's O(N^2 * 27) search
(src/getting_nbd.f90:46-82, src/search_nbd.f90:34-58), and directly against
that search:

  - every example input directory of the reference at its own size (lists
    and counts identical to the host builder) and, at up to 4 cells per axis,
    to the oracle's brute-force search;
  - the 16^3-cell pyrochlore (65536 sites, two shells) identical to the host
    builder, and 64 random sites against a brute-force search written here
    (every j, the images {0, +1, -1} per axis);
  - the colouring equal to the host's (ether_colour_lattice) and proper over
    every shell;
  - a list over capacity: ETHER_ERR_LIMIT with the host builder's message.
"""
import itertools
import os

import numpy as np
import pytest

from ether_amd import engine as E
from ether_amd import workloads
from ether_amd.inputs import read_directory
from tests.helpers import say
from oracle import oracle as o

pytestmark = pytest.mark.gpu

EX = os.path.join(os.path.dirname(__file__), "golden", "examples")


def _args(m, sc=None):
    return (m.structure.abc, m.basis_xyz, m.sc if sc is None else sc, [b == "o" for b in m.tags.bc],
            m.shells, m.tags.nbdfc)


def _same(a, b):
    ca, ia = a
    cb, ib = b
    assert np.array_equal(ca, cb)
    for s in range(ca.shape[0]):
        for i in range(ca.shape[1]):
            assert np.array_equal(ia[s, i, : ca[s, i]], ib[s, i, : cb[s, i]]), (s, i)


def _proper(col, cnt, idx):
    for s in range(cnt.shape[0]):
        mask = np.arange(idx.shape[2])[None, :] < cnt[s][:, None]
        nb = np.where(mask, idx[s], 0)
        clash = (col[nb] == col[:, None]) & mask
        assert not clash.any(), s


@pytest.mark.parametrize("name", ["SimpleCubic_FM", "SimpleCubic_AFM", "SimpleCubic_333", "SimpleCubic_141414",
                                  "Triangular", "Pyrochlore"])
def test_device_builder_on_the_reference_examples(name):
    say(f"device neighbours: example {name}")
    m = read_directory(os.path.join(EX, name))
    cap = m.tags.nbd_capacity
    cnt, idx, col = E.build_neighbours_device(*_args(m), cap)
    _same((cnt, idx), E.build_neighbours(*_args(m), cap))
    hc = E.colour_lattice(*_args(m))
    assert (col is None) == (hc is None)
    if col is not None:
        assert np.array_equal(col, hc)
        _proper(col, cnt, idx)
    sc = tuple(min(x, 4) for x in m.sc)
    cs, isd, _ = E.build_neighbours_device(*_args(m, sc), cap)
    _same((cs, isd), o.neighbours_brute(*_args(m, sc), cap))


def test_device_builder_on_the_pyrochlore_workload():
    say("device neighbours: pyrochlore 16^3 x 16")
    m = workloads.model("pyro_pt_16")
    cnt, idx, col = E.build_neighbours_device(*_args(m), 20)
    _same((cnt, idx), E.build_neighbours(*_args(m), 20))
    assert np.all(cnt[0] == 6) and np.all(cnt[1] == 12)
    assert col is not None and col.max() + 1 == 4
    assert np.array_equal(col, E.colour_lattice(*_args(m)))
    _proper(col, cnt, idx)
    # brute force on 64 sites: every j, every image of the supercell
    abc = np.asarray(m.structure.abc)
    L = np.array(m.sc)
    pos = m.positions
    rng = np.random.default_rng(3)
    shifts = np.array(list(itertools.product((0, 1, -1), repeat=3))) * L
    img = shifts @ abc                                 # (27, 3)
    for i in rng.choice(m.n_sites, 64, replace=False):
        d = np.linalg.norm(pos[i][None, None, :] - (pos[:, None, :] + img[None, :, :]), axis=2)   # (N, 27)
        for s, ds in enumerate(m.shells):
            hit = np.flatnonzero(np.any(np.abs(ds - d) <= m.tags.nbdfc, axis=1))
            hit = hit[hit != i]
            assert np.array_equal(idx[s, i, : cnt[s, i]], hit), (s, i)


def test_device_builder_capacity_error():
    say("device neighbours: capacity")
    with pytest.raises(E.EtherError) as ei:
        E.build_neighbours_device(np.eye(3), [[0, 0, 0]], (4, 4, 4), (0, 0, 0), [1.0], 1e-5, 4)
    assert ei.value.code == E.ETHER_ERR_LIMIT
    assert "site 1 shell 1 has 6 neighbours > nbd_capacity 4" in str(ei.value)
