"""The batched uneven workload: the generator reproduces the committed list of
256 lattices, and every lattice parses and builds with the expected size."""
import json
import os

import numpy as np

from ether_amd import workloads

HERE = os.path.dirname(os.path.abspath(__file__))


def test_generator_reproduces_fixture():
    with open(os.path.join(HERE, "golden", "c5_lattices.json")) as f:
        ref = json.load(f)
    got = workloads.c5_lattices(ref["seed"])
    assert len(got) == 256
    for a, b in zip(got, ref["lattices"]):
        assert a["kind"] == b["kind"] and a["size"] == b["size"] and a["R"] == b["R"]
        assert a["seed"] == b["seed"]
        assert np.isclose(a["J"], b["J"], rtol=0, atol=1e-15)
        assert np.isclose(a["t"], b["t"], rtol=0, atol=1e-15)


def test_ranges_and_models():
    specs = workloads.c5_lattices()
    kinds = {s["kind"] for s in specs}
    assert kinds == set(workloads.C5_KINDS)
    for s in specs:
        assert 8 <= s["R"] <= 32
        assert 0.5 <= abs(s["J"]) <= 1.5 and 0.2 <= s["t"] <= 2.0
        lo, hi = {"sc_ising": (8, 48), "sc_heis": (8, 48), "tri": (24, 128), "pyro": (4, 12)}[s["kind"]]
        assert lo <= s["size"] <= hi
    for s in specs[:12]:
        m = workloads.c5_model(s)
        L = s["size"]
        n = {"sc_ising": L ** 3, "sc_heis": L ** 3, "tri": L * L, "pyro": 16 * L ** 3}[s["kind"]]
        assert m.n_sites == n
        assert m.tags.ising == (s["kind"] == "sc_ising")
        T = workloads.c5_temperatures(s)
        assert len(T) == s["R"] and np.isclose(T[0], 2 * T[-1]) and np.all(np.diff(T) < 0)
