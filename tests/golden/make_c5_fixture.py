#!/usr/bin/env python3
"""Writes tests/golden/c5_lattices.json: the 256 lattices of the batched
uneven workload as drawn by ether_amd.workloads.c5_lattices(seed=1992), so
that a change of the generator (or of numpy's PCG64 stream) is caught."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ether_amd.workloads import c5_lattices  # noqa: E402

if __name__ == "__main__":
    with open(os.path.join(HERE, "c5_lattices.json"), "w") as f:
        json.dump({"seed": 1992, "lattices": c5_lattices(1992)}, f, indent=0)
