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
Every one of them can be shrunk (sc, mcs, repeat, temperatures) for tests.
"""
from __future__ import annotations

import math
from typing import Dict, Optional

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
