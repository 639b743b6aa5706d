"""ETHER input files: in.ether tags, structure.vasp, j_exchange,
single_ion_anisotropy and staggered, and the model arrays built from them.

The reference reads these in Fortran (src/read_input.f90, src/defaults.f90,
src/process_file.f90, src/read_structure.f90, src/j_values.f90,
src/get_sia_values.f90, src/get_staggered_info.f90, src/parameters.f90,
src/generate_supercell.f90).  This module keeps the same tags, defaults,
scalings and failure conditions; a failure raises EtherInputError with the
reference's message where the reference terminates.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

KB_EV = 8.617333262e-5      # src/init.f90:47
MUB_EV = 5.7883818060e-5    # src/parameters.f90:41


class EtherInputError(ValueError):
    """The reference would terminate (src/terminate.f90:20-35)."""


def fortran_nint(x: float) -> int:
    """Fortran NINT: round half away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def _lu(sym: str) -> str:
    """First letter upper, second lower (src/lu.f90 as used on species names)."""
    s = sym.strip()
    if not s:
        return s
    return s[0].upper() + s[1:2].lower()


def _logical(tok: str) -> bool:
    t = tok.strip().lower().lstrip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise EtherInputError(f"cannot read logical from '{tok}'")


def _toks(value: str) -> List[str]:
    # Fortran list-directed input: blanks and commas separate values
    return [t for t in re.split(r"[\s,]+", value.strip()) if t]


@dataclass
class InEther:
    """Tags of in.ether with the defaults of src/defaults.f90:25-56."""
    model: str = "ising"
    tmcs: int = 5000
    tmcs_eq: int = 3000
    ht: float = 100.0
    lt: float = 10.0
    tint: float = 5.0
    spin: Optional[List[float]] = None
    species: Optional[List[str]] = None
    sc: Tuple[int, int, int] = (2, 2, 2)
    stg: bool = False
    bc: Tuple[str, str, str] = ("c", "c", "c")
    repeat: int = 10
    angle: bool = False
    dphi: float = 0.0           # radians (src/read_input.f90:95-97)
    coa: float = 10.0
    zeeman: bool = False
    h: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    g_factor: float = 2.0
    sia: bool = False
    para: bool = False
    j_para: float = 1.0
    ovrr: bool = False
    ovrr_para: float = 0.0
    ovrr_mcs: int = 0
    seed: int = 1992
    nbdfc: float = 1e-5
    pt: bool = False
    exchange_interval: int = 11
    cb: bool = False
    ss_dis: float = 4.0
    ss_lat: float = 2.0
    ss_direc: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    ssp: bool = False
    spiral_capacity: int = 5
    nbd_capacity: int = 20

    @property
    def ising(self) -> bool:
        return self.model == "ising"

    @property
    def xyz(self) -> bool:
        return self.model == "xyz"

    def temperatures(self) -> np.ndarray:
        """src/ether.f90:79-84: nscan = nint((ht-lt)/tint)+1, T_k = ht - tint*(k-1)."""
        nscan = fortran_nint((self.ht - self.lt) / self.tint) + 1
        return np.array([self.ht - self.tint * k for k in range(nscan)], np.float64)


def process_lines(text: str) -> List[str]:
    """src/process_file.f90:45-69: tabs removed, '!' or '#' starts a comment,
    empty lines dropped."""
    out = []
    for line in text.splitlines():
        line = line.replace("\t", "")
        cut = len(line)
        for ch in "!#":
            p = line.find(ch)
            if p >= 0:
                cut = min(cut, p)
        line = line[:cut]
        if line.strip():
            out.append(line)
    return out


def parse_in_ether(text: str, n_species: int) -> InEther:
    """src/read_input.f90:41-150."""
    t = InEther()
    for raw in process_lines(text):
        line = raw.lower()
        j = line.find("=")
        if j < 0:
            raise EtherInputError(f"Missing '=' in the given line: {raw.strip()}")
        key = line[:j].strip()
        v = _toks(line[j + 1:])
        try:
            if key == "model":
                t.model = v[0][:5]
                if t.model not in ("ising", "xyz"):
                    raise EtherInputError(f"unknown model '{v[0]}'")
            elif key == "mcs":
                t.tmcs, t.tmcs_eq = int(v[0]), int(v[1])
                if t.tmcs <= t.tmcs_eq:
                    raise EtherInputError("TOTAL MCS  < EQUILIBRATIONS MCS steps")
            elif key == "temp":
                t.ht, t.lt, t.tint = float(v[0]), float(v[1]), float(v[2])
            elif key == "spin":
                t.spin = [float(x) for x in v[:n_species]]
                if len(t.spin) < n_species:
                    raise EtherInputError("'spin' needs one value per species of structure.vasp")
            elif key == "species":
                t.species = [_lu(x) for x in v]
            elif key == "sc":
                t.sc = (int(v[0]), int(v[1]), int(v[2]))
            elif key == "stg":
                t.stg = _logical(v[0])
            elif key == "bc":
                t.bc = (v[0][0], v[1][0], v[2][0])
            elif key == "repeat":
                t.repeat = int(v[0])
            elif key == "angle":
                t.angle = _logical(v[0])
                t.dphi = math.pi / 180.0 * float(v[1]) if len(v) > 1 else 0.0
            elif key == "coa":
                t.coa = float(v[0])
            elif key == "zeeman":
                t.zeeman = _logical(v[0])
                t.h = (float(v[1]), float(v[2]), float(v[3]))
            elif key == "g_factor":
                t.g_factor = float(v[0])
            elif key == "sia":
                t.sia = _logical(v[0])
            elif key == "para":
                t.para = _logical(v[0])
                t.j_para = abs(float(v[1]))
                if t.j_para == 0.0 and t.para:
                    raise EtherInputError("parameter value is set ON and value cannot be zero")
            elif key == "ovrr":
                t.ovrr = _logical(v[0])
                t.ovrr_para = abs(float(v[1]))
                t.ovrr_mcs = int(float(v[2]))
            elif key == "seed":
                t.seed = int(v[0])
            elif key == "nbdfc":
                t.nbdfc = float(v[0])
            elif key == "pt":
                t.pt = _logical(v[0])
                t.exchange_interval = int(v[1])
            elif key == "cb":
                t.cb = _logical(v[0])
            elif key == "ss_dis":
                t.ss_dis = float(v[0])
            elif key == "ss_lat":
                t.ss_lat = float(v[0])
            elif key == "ss_direc":
                t.ss_direc = (float(v[0]), float(v[1]), float(v[2]))
            elif key == "ssp":
                t.ssp = _logical(v[0])
            elif key == "spiral_capacity":
                t.spiral_capacity = int(v[0])
            elif key == "nbd_capacity":
                t.nbd_capacity = int(v[0])
            else:
                raise EtherInputError(f"Found unknown/missing information: {raw.strip()}")
        except (IndexError, ValueError) as e:
            if isinstance(e, EtherInputError):
                raise
            raise EtherInputError(f"cannot read tag '{key}' from: {raw.strip()}") from e
    return t


@dataclass
class Structure:
    """structure.vasp (src/read_structure.f90:51-106)."""
    title: str
    abc: np.ndarray            # 3x3, rows are lattice vectors (Angstrom)
    species: List[str]         # species names, file order
    counts: List[int]
    cart: np.ndarray           # (n_atoms, 3) Cartesian positions, file order

    @property
    def n_species(self) -> int:
        return len(self.species)


def parse_structure(text: str) -> Structure:
    lines = [ln for ln in text.splitlines()]
    if len(lines) < 8:
        raise EtherInputError("Incorrect structure file")
    title = lines[0].strip()
    abc = np.array([[float(x) for x in lines[2 + i].split()[:3]] for i in range(3)], np.float64)
    species = [_lu(s) for s in lines[5].split()]
    counts = [int(x) for x in lines[6].split()[: len(species)]]
    label = lines[7].strip().lower()
    n = sum(counts)
    coords = np.array([[float(x) for x in lines[8 + i].split()[:3]] for i in range(n)], np.float64)
    if label.startswith("direct"):
        # r_cart(i) = sum_j abc(j,i) * frac(j)  (src/frac_to_cart.f90:45-49)
        cart = coords @ abc
    elif label.startswith("cartesian"):
        cart = coords
    else:
        raise EtherInputError("Incorrect structure file")
    return Structure(title, abc, species, counts, cart)


@dataclass
class Exchange:
    """j_exchange (src/j_values.f90:39-87), raw meV values before scaling."""
    shells: List[float]
    pairs: List[List[Tuple[str, str, Tuple[float, float, float]]]]


def parse_j_exchange(text: str) -> Exchange:
    toks_lines = [ln.split() for ln in text.splitlines()]
    it = iter(toks_lines)

    def nxt():
        for tl in it:
            if tl:
                return tl
        raise EtherInputError("j_exchange ended early")

    n_shells = int(nxt()[0])
    shells, pairs = [], []
    for _ in range(n_shells):
        tl = nxt()
        d, npairs = float(tl[0]), int(tl[1])
        shells.append(d)
        ps = []
        for _ in range(npairs):
            tl = nxt()
            ps.append((_lu(tl[0]), _lu(tl[1]), (float(tl[2]), float(tl[3]), float(tl[4]))))
        pairs.append(ps)
    return Exchange(shells, pairs)


def parse_sia(text: str) -> List[Tuple[str, Tuple[float, float, float]]]:
    """single_ion_anisotropy: `sp Dx Dy Dz` lines (src/get_sia_values.f90:44-61)."""
    out = []
    for ln in text.splitlines():
        tl = ln.split()
        if len(tl) >= 4:
            out.append((_lu(tl[0]), (float(tl[1]), float(tl[2]), float(tl[3]))))
    return out


def parse_staggered(text: str) -> List[Tuple[float, str]]:
    """staggered: `c_i sp` lines (src/get_staggered_info.f90:59-74)."""
    out = []
    for ln in text.splitlines():
        tl = ln.split()
        if len(tl) >= 2:
            out.append((float(tl[0]), _lu(tl[1])))
    return out


@dataclass
class Model:
    """Everything the engine needs, in the reference's units and scalings."""
    tags: InEther
    structure: Structure
    n_species: int
    basis_species: np.ndarray        # (n_basis,) 1-based species id per basis site
    basis_xyz: np.ndarray            # (n_basis, 3)
    basis_file_index: np.ndarray     # (n_basis,) 1-based atom index in structure.vasp
    sc: Tuple[int, int, int]
    n_sites: int
    species: np.ndarray              # (N,) 1-based
    stagger: np.ndarray              # (N,)
    positions: np.ndarray            # (N, 3)
    shells: np.ndarray               # (n_shells,)
    j_exc: np.ndarray                # (n_shells, nsp, nsp, 3) scaled
    spin: np.ndarray                 # (nsp,)
    sia_vec: Optional[np.ndarray]    # (nsp, 3) scaled, or None
    h: np.ndarray                    # field (after PARA scaling)
    g_factor: float
    mu_b: float
    kb: float
    nn_count: Optional[np.ndarray] = None
    nn_idx: Optional[np.ndarray] = None
    colour: Optional[np.ndarray] = None   # periodic colouring of the neighbour graph (None: greedy at commit)
    initial_spins: Optional[np.ndarray] = None  # (N, 3) generate_supercell start state
    initial_phi: Optional[np.ndarray] = None
    gpu: "GpuTags" = field(default_factory=lambda: GpuTags())

    @property
    def model_id(self) -> int:
        return 0 if self.tags.ising else 1

    def betas(self, temps) -> np.ndarray:
        return 1.0 / (self.kb * np.asarray(temps, np.float64))


def build_model(tags: InEther, st: Structure, ex: Exchange, sia_lines=None, stg_lines=None) -> Model:
    """generate_supercell + j_values + get_sia_values + parameters."""
    nsp = st.n_species
    if tags.spin is None:
        tags.spin = [1.0] * nsp
    if tags.species is None:
        tags.species = list(st.species)
    # basis filtering in the order of the `species` tag (src/generate_supercell.f90:65-101)
    starts = np.concatenate([[0], np.cumsum(st.counts)])
    b_sp, b_xyz, b_idx = [], [], []
    for name in tags.species:
        for j, sname in enumerate(st.species):
            if name == sname:
                for k in range(starts[j], starts[j + 1]):
                    b_sp.append(j + 1)
                    b_xyz.append(st.cart[k])
                    b_idx.append(k + 1)
    if not b_sp:
        raise EtherInputError("Selected ions from 'in.ether' are not present in 'structure.vasp' file")
    b_sp = np.array(b_sp, np.int32)
    b_xyz = np.array(b_xyz, np.float64)
    b_idx = np.array(b_idx, np.int32)
    nb = len(b_sp)
    sc = tuple(int(x) for x in tags.sc)
    n = sc[0] * sc[1] * sc[2] * nb
    # supercell: k (z) outer, j, i, basis l inner (src/generate_supercell.f90:115-132)
    kk, jj, ii, ll = np.meshgrid(np.arange(sc[2]), np.arange(sc[1]), np.arange(sc[0]),
                                 np.arange(nb), indexing="ij")
    kk, jj, ii, ll = kk.ravel(), jj.ravel(), ii.ravel(), ll.ravel()
    pos = (st.abc[0][None, :] * ii[:, None] + st.abc[1][None, :] * jj[:, None] +
           st.abc[2][None, :] * kk[:, None] + b_xyz[ll])
    species = b_sp[ll]
    # staggered signs (src/generate_supercell.f90:127, src/get_staggered_info.f90:42-82)
    if tags.cb:
        stagger = np.where(((ii + 1) + (jj + 1) + (kk + 1) + (ll + 1)) % 2 == 0, 1.0, -1.0)
    else:
        stg_ids = np.ones(nsp, np.float64)
        if tags.stg and stg_lines is not None:
            for c, name in stg_lines:
                for k2, sname in enumerate(st.species):
                    if name == sname:
                        stg_ids[k2] = c
        stagger = stg_ids[species - 1]
    spin = np.array(tags.spin[:nsp], np.float64)
    # j_values (src/j_values.f90:64-85): J * s_a * s_b, symmetric in (a, b)
    ns = len(ex.shells)
    j_exc = np.zeros((ns, nsp, nsp, 3), np.float64)
    name_to_id = {s: k for k, s in enumerate(st.species)}
    for sh, ps in enumerate(ex.pairs):
        for a, b, J in ps:
            if a not in name_to_id or b not in name_to_id:
                raise EtherInputError(f"j_exchange names species '{a}'/'{b}' not in structure.vasp")
            ia, ib = name_to_id[a], name_to_id[b]
            v = np.array(J, np.float64) * spin[ia] * spin[ib]
            j_exc[sh, ia, ib] = v
            j_exc[sh, ib, ia] = v
    # single-ion anisotropy (src/get_sia_values.f90:32-62): D * s^2, XYZ only
    sia_vec = None
    if tags.sia and tags.xyz:
        if sia_lines is None:
            raise EtherInputError("SIA is .TRUE. but file 'single_ion_anisotropy' is not found")
        sia_vec = np.zeros((nsp, 3), np.float64)
        for name, D in sia_lines:
            if name in name_to_id:
                k2 = name_to_id[name]
                sia_vec[k2] = np.array(D) * spin[k2] ** 2
    h = np.array(tags.h, np.float64)
    # parameters (src/parameters.f90:24-58)
    if tags.para:
        j_exc = j_exc / tags.j_para
        if sia_vec is not None:
            sia_vec = sia_vec / tags.j_para
        kb, mub, g = 1.0, 1.0, 1.0
        if tags.zeeman:
            h = h / tags.j_para
    else:
        j_exc = j_exc / 1000.0
        if sia_vec is not None:
            sia_vec = sia_vec / 1000.0
        kb, mub, g = KB_EV, MUB_EV, tags.g_factor
    # start state (src/generate_supercell.f90:34-63): Ising (-1)^i by file index,
    # XYZ random per basis site (a pure function of SEED, drawn by the library
    # so that every driver starts from the same state), copied to every cell
    from .engine import start_state
    init, phi0 = start_state(tags.seed, 0 if tags.ising else 1, 1 if (tags.angle and not tags.ising) else 0,
                             b_idx)
    return Model(tags=tags, structure=st, n_species=nsp, basis_species=b_sp, basis_xyz=b_xyz,
                 basis_file_index=b_idx, sc=sc, n_sites=n, species=species.astype(np.int32),
                 stagger=stagger.astype(np.float64), positions=pos, shells=np.array(ex.shells),
                 j_exc=j_exc, spin=spin, sia_vec=sia_vec, h=h, g_factor=g, mu_b=mub, kb=kb,
                 initial_spins=init[ll], initial_phi=phi0[ll])


GPU_FILE = "gpu.ether"
BATCH_FILE = "batch.ether"


@dataclass
class GpuTags:
    """Engine settings of a run directory's gpu.ether: the tags the reference
    has no counterpart for.  They live in their own file because the
    reference's parser stops on any unknown key (src/read_input.f90:144-146),
    so in.ether stays a valid reference input.  Same syntax as in.ether
    (key = values; '!' or '#' comments; Fortran logicals)."""
    ladder: str = "linear"      # linear | learned (tempering runs)
    train_mcs: int = 0          # learned: MCS of the training windows (0: tmcs_eq // 2)
    windows: int = 4            # learned: training windows
    autotune: int = 0           # sweeps timed per launch shape (0: off)
    spw: int = 0                # launch: sites per wavefront (0: engine default)
    xcd: int = 1                # launch: XCD-aware block order
    lanes: int = 0              # lanes per site of the sweep (0: chosen from the replica count)
    graphs: bool = True         # replay captured HIP graphs
    compat: int = 0             # compat flags (1: PT on E_tot, 2: OVRR as a separate pass,
                                # 4: measurement as a separate pass, 8: repeats one after
                                # another, each from the previous one's final state)
    chunk: int = 0              # MCS per ether_run call (0: the whole run)


def parse_gpu_ether(text: str) -> GpuTags:
    g = GpuTags()
    for raw in process_lines(text):
        line = raw.lower()
        j = line.find("=")
        if j < 0:
            raise EtherInputError(f"Missing '=' in the given line: {raw.strip()}")
        key = line[:j].strip()
        v = _toks(line[j + 1:])
        try:
            if key == "ladder":
                g.ladder = v[0]
                if g.ladder not in ("linear", "learned"):
                    raise EtherInputError(f"ladder must be 'linear' or 'learned', not '{v[0]}'")
                if len(v) > 1:
                    g.train_mcs = int(v[1])
                if len(v) > 2:
                    g.windows = int(v[2])
                if g.train_mcs < 0 or g.windows < 1:
                    raise EtherInputError("ladder: training MCS must be >= 0 and windows >= 1")
            elif key == "autotune":
                g.autotune = int(v[0])
            elif key == "launch":
                g.spw, g.xcd = int(v[0]), int(v[1])
                if g.spw not in (1, 2, 4, 8, 16) or g.xcd not in (0, 1, 2):
                    raise EtherInputError("launch: sites per wave in {1,2,4,8,16}, XCD map 0, 1 or 2")
            elif key == "lanes":
                g.lanes = int(v[0])
                if not 0 <= g.lanes <= 64:
                    raise EtherInputError("lanes: 0 (automatic) or 1 to 64 lanes per site (Ising: a power "
                                          "of two from 8)")
            elif key == "graphs":
                g.graphs = _logical(v[0])
            elif key == "compat":
                g.compat = int(v[0])
                if g.compat < 0 or g.compat > 15:
                    raise EtherInputError("compat: a sum of the flags 1, 2, 4 and 8")
            elif key == "chunk":
                g.chunk = int(v[0])
                if g.chunk < 0:
                    raise EtherInputError("chunk must be >= 0")
            else:
                raise EtherInputError(f"Found unknown/missing information in {GPU_FILE}: {raw.strip()}")
        except (IndexError, ValueError) as e:
            if isinstance(e, EtherInputError):
                raise
            raise EtherInputError(f"cannot read tag '{key}' from: {raw.strip()}") from e
    return g


def parse_batch_manifest(text: str, base: str = ".") -> List[str]:
    """batch.ether: one run directory per line (relative to the manifest's
    directory), comments as in in.ether.  The runs advance together through
    one schedule per group of equal (mcs, coa, ovrr, pt) tags."""
    out = []
    for raw in process_lines(text):
        d = raw.strip()
        out.append(d if os.path.isabs(d) else os.path.normpath(os.path.join(base, d)))
    if not out:
        raise EtherInputError(f"{BATCH_FILE} lists no run directory")
    if len(set(out)) != len(out):
        raise EtherInputError(f"{BATCH_FILE} lists a run directory twice")
    return out


def read_directory(path: str) -> Model:
    """Read in.ether and the input files of a run directory, as the reference does
    from its current directory."""
    def rd(name):
        with open(os.path.join(path, name)) as f:
            return f.read()

    if not os.path.exists(os.path.join(path, "structure.vasp")):
        raise EtherInputError("Required file 'structure.vasp' is not present")
    st = parse_structure(rd("structure.vasp"))
    if not os.path.exists(os.path.join(path, "in.ether")):
        raise EtherInputError("Required file 'in.ether' is not present")
    tags = parse_in_ether(rd("in.ether"), st.n_species)
    if not os.path.exists(os.path.join(path, "j_exchange")):
        raise EtherInputError("Required file 'j_exchange' is not present")
    ex = parse_j_exchange(rd("j_exchange"))
    sia = parse_sia(rd("single_ion_anisotropy")) if (
        tags.sia and os.path.exists(os.path.join(path, "single_ion_anisotropy"))) else None
    stg = parse_staggered(rd("staggered")) if (
        tags.stg and os.path.exists(os.path.join(path, "staggered"))) else None
    m = build_model(tags, st, ex, sia, stg)
    if os.path.exists(os.path.join(path, GPU_FILE)):
        m.gpu = parse_gpu_ether(rd(GPU_FILE))
    return m
