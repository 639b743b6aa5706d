"""The five benchmark workloads, written as ETHER input files and parsed by the
same reader as a user's run directory.

  sc_ising_16       Ising simple cubic 16^3, FM J = 1 meV, T 100 -> 10 K step 5,
                    MCS 5000/3000, coa 10, REPEAT 10, seed 1992
  sc_heis_afm_32    Heisenberg simple cubic 32^3, AFM J = 1 meV, Marsaglia moves,
                    OVRR every MCS, T 60 -> 6 K (19), MCS 5000/3000, REPEAT 10
  tri_afm_96        triangular 96x96x1 Heisenberg AFM, J = 1 meV, easy-axis SIA
                    D_z = -0.1 meV, H_z = 1 T, g = 2, ANGLE (5 deg) or Marsaglia,
                    OVRR every 5 MCS, 32 T from 20 -> 0.5 K
  pyro_pt_16        pyrochlore 16^3 cubic cells x 16 sites, J1 = 1, J2 = 0.05 meV,
                    tempering with 64 replicas, swaps every 11 MCS, T 1 -> 0.05 K,
                    MCS 20000/5000, REPEAT 4
  c5 (c5_lattices)  256 uneven lattices for one batch: simple-cubic Ising and
                    Heisenberg (edge 8..48), triangular (24..128 squared),
                    pyrochlore (4..12 cubed cells), J = +-U(0.5, 1.5) meV, each
                    with its own ladder of 8..32 replicas from T0 to 2 T0,
                    T0 = U(0.2, 2) |J| / kB
Every one of them can be shrunk (sc, mcs, repeat, temperatures) for tests.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional

import numpy as np

from .inputs import Model, build_model, parse_in_ether, parse_j_exchange, parse_sia, parse_structure

SC_STRUCTURE = """simple cubic
1.0
 1.0 0.0 0.0
 0.0 1.0 0.0
 0.0 0.0 1.0
 A
 1
Cartesian
 0.0 0.0 0.0
"""

TRI_STRUCTURE = """triangular
1.0
 1.0 0.0 0.0
 0.5 0.8660254037844386 0.0
 0.0 0.0 10.0
 A
 1
Cartesian
 0.0 0.0 0.0
"""

PYRO_A = 8.4072999954
PYRO_STRUCTURE = f"""Pyrochlore
1.0
 {PYRO_A:.10f} 0.0 0.0
 0.0 {PYRO_A:.10f} 0.0
 0.0 0.0 {PYRO_A:.10f}
 A
 16
Direct
 0.000 0.000 0.000
 0.750 0.250 0.500
 0.250 0.750 0.500
 0.250 0.500 0.750
 0.750 0.500 0.250
 0.500 0.750 0.250
 0.500 0.250 0.750
 0.000 0.500 0.500
 0.750 0.750 0.000
 0.250 0.250 0.000
 0.250 0.000 0.250
 0.750 0.000 0.750
 0.500 0.000 0.500
 0.000 0.750 0.750
 0.000 0.250 0.250
 0.500 0.500 0.000
"""

PYRO_D1 = PYRO_A * math.sqrt(2.0) / 4.0
PYRO_D2 = PYRO_A * math.sqrt(6.0) / 4.0


def _tags(d: Dict[str, str]) -> str:
    return "\n".join(f"{k} = {v}" for k, v in d.items()) + "\n"


def files(name: str, **over) -> Dict[str, str]:
    """Input files of a workload; keyword overrides replace in.ether tags."""
    if name == "sc_ising_16":
        tags = dict(model="ising", mcs="5000 3000", temp="100 10 5", spin="1", species="A",
                    sc="16 16 16", bc="c c c", repeat="10", coa="10", seed="1992")
        out = dict(structure=SC_STRUCTURE, j_exchange="1\n1.0 1\nA A -1 -1 -1\n")
    elif name == "sc_heis_afm_32":
        tags = dict(model="xyz", mcs="5000 3000", temp="60 6 3", spin="1", species="A",
                    sc="32 32 32", bc="c c c", repeat="10", coa="10", seed="1992",
                    angle=".F. 0", ovrr=".T. 1.0 1", cb=".T.")
        out = dict(structure=SC_STRUCTURE, j_exchange="1\n1.0 1\nA A 1 1 1\n")
    elif name == "tri_afm_96":
        tags = dict(model="xyz", mcs="5000 3000", temp="20 0.5 0.629032", spin="1", species="A",
                    sc="96 96 1", bc="c c c", repeat="10", coa="10", seed="1992",
                    angle=".T. 5", zeeman=".T. 0 0 1", g_factor="2", sia=".T.",
                    ovrr=".T. 1.0 5")
        out = dict(structure=TRI_STRUCTURE, j_exchange="1\n1.0 1\nA A 1 1 1\n",
                   single_ion_anisotropy="A 0 0 -0.1\n")
    elif name == "pyro_pt_16":
        tags = dict(model="xyz", mcs="20000 5000", temp="1 0.05 0.0150793650793651", spin="1",
                    species="A", sc="16 16 16", bc="c c c", repeat="4", coa="10", seed="1992",
                    angle=".F. 0", pt=".T. 11")
        out = dict(structure=PYRO_STRUCTURE,
                   j_exchange=f"2\n{PYRO_D1:.6f} 1\nA A 1 1 1\n{PYRO_D2:.6f} 1\nA A 0.05 0.05 0.05\n")
    else:
        raise KeyError(name)
    for k, v in over.items():
        tags[k] = v
    out["in.ether"] = _tags(tags)
    return out


def model(name: str, nbdfc: Optional[float] = None, **over) -> Model:
    f = files(name, **over)
    st = parse_structure(f["structure"])
    t = parse_in_ether(f["in.ether"], st.n_species)
    if nbdfc is not None:
        t.nbdfc = nbdfc
    ex = parse_j_exchange(f["j_exchange"])
    sia = parse_sia(f["single_ion_anisotropy"]) if "single_ion_anisotropy" in f else None
    return build_model(t, st, ex, sia, None)


def write_dir(name: str, path: str, **over) -> None:
    """Write the workload as a run directory (in.ether, structure.vasp, ...)."""
    import os
    f = files(name, **over)
    os.makedirs(path, exist_ok=True)
    names = {"structure": "structure.vasp", "j_exchange": "j_exchange", "in.ether": "in.ether",
             "single_ion_anisotropy": "single_ion_anisotropy"}
    for k, v in f.items():
        with open(os.path.join(path, names[k]), "w") as fh:
            fh.write(v)


# ---------------------------------------------------------------------------
# The batched uneven workload
# ---------------------------------------------------------------------------
KB_MEV_PER_K = 0.08617333262
C5_KINDS = ("sc_ising", "sc_heis", "tri", "pyro")


def c5_lattices(seed: int = 1992, n: int = 256) -> List[dict]:
    """The lattices of the batched workload, drawn with numpy's PCG64 seeded
    with `seed`, in this order per lattice: kind (uniform over C5_KINDS), size
    (simple-cubic edge U{8..48}, triangular edge U{24..128}, pyrochlore cells
    per edge U{4..12}), |J| ~ U(0.5, 1.5) meV with a uniform sign (negative =
    ferromagnetic in ETHER's E = sum J S.S), t ~ U(0.2, 2), replicas
    R ~ U{8..32}.  The ladder runs geometrically from T0 = t |J| / kB to 2 T0."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        kind = C5_KINDS[int(rng.integers(4))]
        if kind in ("sc_ising", "sc_heis"):
            size = int(rng.integers(8, 49))
        elif kind == "tri":
            size = int(rng.integers(24, 129))
        else:
            size = int(rng.integers(4, 13))
        J = float(rng.uniform(0.5, 1.5)) * (1.0 if rng.random() < 0.5 else -1.0)
        t = float(rng.uniform(0.2, 2.0))
        R = int(rng.integers(8, 33))
        out.append(dict(kind=kind, size=size, J=J, t=t, R=R, seed=seed * 1000 + k))
    return out


def c5_files(spec: dict) -> Dict[str, str]:
    """Input files of one lattice of the batched workload."""
    kind, L, J = spec["kind"], spec["size"], spec["J"]
    tags = dict(model="ising" if kind == "sc_ising" else "xyz", mcs="1000 100", temp="1 1 1",
                spin="1", species="A", bc="c c c", repeat="1", coa="10", seed=str(spec["seed"]))
    if kind != "sc_ising":
        tags["angle"] = ".F. 0"
    if kind in ("sc_ising", "sc_heis"):
        tags["sc"] = f"{L} {L} {L}"
        files_ = dict(structure=SC_STRUCTURE, j_exchange=f"1\n1.0 1\nA A {J} {J} {J}\n")
    elif kind == "tri":
        tags["sc"] = f"{L} {L} 1"
        files_ = dict(structure=TRI_STRUCTURE, j_exchange=f"1\n1.0 1\nA A {J} {J} {J}\n")
    else:
        tags["sc"] = f"{L} {L} {L}"
        files_ = dict(structure=PYRO_STRUCTURE, j_exchange=f"1\n{PYRO_D1:.6f} 1\nA A {J} {J} {J}\n")
    files_["in.ether"] = _tags(tags)
    return files_


def c5_model(spec: dict) -> Model:
    f = c5_files(spec)
    st = parse_structure(f["structure"])
    t = parse_in_ether(f["in.ether"], st.n_species)
    return build_model(t, st, parse_j_exchange(f["j_exchange"]), None, None)


def c5_temperatures(spec: dict) -> np.ndarray:
    """The lattice's ladder in K, hottest first."""
    t0 = spec["t"] * abs(spec["J"]) / KB_MEV_PER_K
    return np.geomspace(2.0 * t0, t0, spec["R"])
