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
