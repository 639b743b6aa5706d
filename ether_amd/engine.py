"""ctypes binding of the engine's C-ABI (include/ether_gpu.h).

This is the Python face of the same boundary the Fortran driver binds through
``iso_c_binding``.  Every call goes to ``ether_amd/lib/libether_gpu.so``; there
is no CPU fallback: a missing library or a missing gfx950 device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

ETHER_OK = 0
ETHER_ERR_ARG = 1
ETHER_ERR_STATE = 2
ETHER_ERR_HIP = 3
ETHER_ERR_LIMIT = 4
ETHER_ERR_NODEV = 5
PLACE_LOG_T = 1
PLACE_LENGTH = 2

MODEL_ISING = 0
MODEL_XYZ = 1
INIT_FRESH = 1
INIT_RANDOM = 2
TIME_SWEEP, TIME_OVRR, TIME_MEASURE, TIME_PT, TIME_SQ, TIME_SWEEP_MEAS, TIME_RESIDENT = 0, 1, 2, 3, 4, 5, 6
COMPAT_PT_ETOT = 1
COMPAT_SPLIT_OVRR = 2
COMPAT_MEASURE_PASS = 4
COMPAT_CHAIN_REPEATS = 8

MOM_N, MOM_E, MOM_E2, MOM_E4, MOM_M, MOM_M2, MOM_M4, MOM_ACC = range(8)
MOM_COUNT = 8
MLP_NPARAM = 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libether_gpu.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

# (name, restype, argtypes); array arguments are passed as void* so that None is allowed
EXPORTS = [
    ("ether_build_neighbours", C.c_int,
     [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, C.c_double, C.c_int32, _vp, _vp]),
    ("ether_build_neighbours_device", C.c_int,
     [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, C.c_double, C.c_int32, _vp, _vp, _vp, _vp]),
    ("ether_colour_lattice", C.c_int,
     [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, C.c_double, _vp, _vp]),
    ("ether_build_spiral_partners", C.c_int,
     [C.c_int32, _vp, _vp, C.c_double, C.c_double, C.c_int32, _vp, _vp]),
    ("ether_start_state", C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp]),
    ("ether_ctx_create", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_vp)]),
    ("ether_ctx_destroy", None, [_vp]),
    ("ether_last_error", C.c_int, [_vp, C.c_char_p, C.c_int32]),
    ("ether_set_lattice", C.c_int, [_vp, C.c_int32, _vp, _vp, C.c_int32, C.c_int32, _vp, _vp, _vp]),
    ("ether_set_exchange", C.c_int, [_vp, _vp, _vp]),
    ("ether_set_zeeman", C.c_int, [_vp, C.c_int32, _vp, C.c_double, C.c_double]),
    ("ether_set_sia", C.c_int, [_vp, C.c_int32, _vp]),
    ("ether_set_proposal", C.c_int, [_vp, C.c_int32, C.c_double]),
    ("ether_set_compat", C.c_int, [_vp, C.c_int32]),
    ("ether_commit", C.c_int, [_vp]),
    ("ether_get_layout", C.c_int, [_vp, _vp, _vp, _vp]),
    ("ether_set_betas", C.c_int, [_vp, _vp]),
    ("ether_retune_betas", C.c_int, [_vp, _vp]),
    ("ether_set_ladders", C.c_int, [_vp, C.c_int32, _vp]),
    ("ether_init_spins", C.c_int, [_vp, C.c_int32]),
    ("ether_sweep", C.c_int, [_vp, C.c_int32, _vp]),
    ("ether_overrelax", C.c_int, [_vp, C.c_double]),
    ("ether_measure", C.c_int, [_vp, C.c_int32, _vp, _vp]),
    ("ether_get_energy_parts", C.c_int, [_vp, _vp, _vp, _vp]),
    ("ether_read_moments", C.c_int, [_vp, _vp]),
    ("ether_reset_moments", C.c_int, [_vp]),
    ("ether_set_spiral", C.c_int, [_vp, C.c_int32, _vp]),
    ("ether_read_spiral", C.c_int, [_vp, _vp, _vp]),
    ("ether_pt_swap", C.c_int, [_vp, C.c_int32, _vp]),
    ("ether_pt_state", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("ether_run", C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                            C.c_int32, _vp]),
    ("ether_set_graphs", C.c_int, [_vp, C.c_int32]),
    ("ether_set_resident", C.c_int, [_vp, C.c_int32]),
    ("ether_set_dft_resident", C.c_int, [_vp, C.c_int32]),
    ("ether_get_resident", C.c_int, [_vp, C.POINTER(C.c_int32)]),
    ("ether_run_prepare", C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    C.c_double, C.c_int32]),
    ("ether_set_launch", C.c_int, [_vp, C.c_int32, C.c_int32]),
    ("ether_set_lanes", C.c_int, [_vp, C.c_int32]),
    ("ether_batch_create", C.c_int, [_vp, C.c_int32, C.POINTER(_vp)]),
    ("ether_batch_run", C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_double, C.c_int32, _vp]),
    ("ether_batch_set_graphs", C.c_int, [_vp, C.c_int32]),
    ("ether_batch_get_resident", C.c_int, [_vp, C.POINTER(C.c_int32)]),
    ("ether_batch_destroy", None, [_vp]),
    ("ether_autotune", C.c_int, [_vp, C.c_int32, _vp, _vp, _vp]),
    ("ether_set_timing", C.c_int, [_vp, C.c_int32]),
    ("ether_get_timing", C.c_int, [_vp, C.c_int32, _vp, _vp]),
    ("ether_synchronize", C.c_int, [_vp]),
    ("ether_get_spins", C.c_int, [_vp, C.c_int32, _vp, _vp]),
    ("ether_set_spins", C.c_int, [_vp, C.c_int32, _vp, _vp]),
    ("ether_local_fields", C.c_int, [_vp, C.c_int32, _vp, _vp]),
    ("ether_structure_factor", C.c_int, [_vp, _vp, _vp]),
    ("ether_structure_factor_basis", C.c_int, [_vp, _vp, C.c_int32, _vp, _vp]),
    ("ether_structure_factor_line", C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                              C.c_int32, _vp]),
    ("ether_structure_factor_line_noalloc", C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                                      C.c_int32, _vp]),
    ("ether_mlp_init", C.c_int, [_vp, C.c_uint64]),
    ("ether_mlp_set_params", C.c_int, [_vp, _vp]),
    ("ether_mlp_get_params", C.c_int, [_vp, _vp]),
    ("ether_mlp_loss_grad", C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, _vp, _vp]),
    ("ether_mlp_forward", C.c_int, [_vp, C.c_int32, _vp, _vp]),
    ("ether_mlp_train", C.c_int, [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, C.c_float, _vp]),
    ("ether_ladder_place", C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_int32, C.c_int32, _vp]),
    ("ether_ladder_train", C.c_int, [_vp, C.c_int32, _vp, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                     C.c_int32, C.c_uint64, _vp, _vp, _vp]),
]

_lib: Optional[C.CDLL] = None


class EtherError(RuntimeError):
    """A nonzero status from the C-ABI, with the library's message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libether_gpu.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EtherError(ETHER_ERR_STATE,
                         f"{path} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, res, args in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _arr(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def _check(rc: int, ctx=None) -> None:
    if rc != ETHER_OK:
        lib = load_library()
        buf = C.create_string_buffer(1024)
        lib.ether_last_error(ctx, buf, 1024)
        raise EtherError(rc, buf.value.decode(errors="replace"))


def start_state(seed, model, angle, file_index):
    """Start state of the basis sites (see ether_start_state): (S [nb, 3], phi [nb])."""
    lib = load_library()
    fi = _arr(file_index, np.int32)
    S = np.zeros((fi.size, 3))
    phi = np.zeros(fi.size)
    _check(lib.ether_start_state(C.c_uint64(int(seed)), int(model), int(angle), fi.size, _ptr(fi),
                                 _ptr(S), _ptr(phi)))
    return S, phi


def build_neighbours_device(abc, basis_xyz, sc, open_bc, shell_d, tol=1e-5, capacity=20):
    """Neighbour lists built on the GPU, and the periodic colouring:
    (nn_count (shells, N), nn_idx (shells, N, capacity), colour (N,) or None)."""
    lib = load_library()
    abc = _arr(abc, np.float64).reshape(3, 3)
    basis = _arr(basis_xyz, np.float64).reshape(-1, 3)
    sc = _arr(sc, np.int32)
    obc = _arr([1 if b else 0 for b in open_bc], np.int32)
    shells = _arr(shell_d, np.float64)
    n = int(np.prod(sc)) * basis.shape[0]
    cnt = np.zeros(len(shells) * n, np.int32)
    idx = np.full(len(shells) * n * capacity, -1, np.int32)
    col = np.zeros(n, np.int32)
    nc = C.c_int32()
    _check(lib.ether_build_neighbours_device(_ptr(abc), basis.shape[0], _ptr(basis), _ptr(sc), _ptr(obc),
                                             len(shells), _ptr(shells), float(tol), int(capacity),
                                             _ptr(cnt), _ptr(idx), _ptr(col), C.byref(nc)))
    return (cnt.reshape(len(shells), n), idx.reshape(len(shells), n, capacity),
            col if nc.value > 0 else None)


def colour_lattice(abc, basis_xyz, sc, open_bc, shell_d, tol=1e-5):
    """The periodic colouring on the host: colour (N,) or None."""
    lib = load_library()
    abc = _arr(abc, np.float64).reshape(3, 3)
    basis = _arr(basis_xyz, np.float64).reshape(-1, 3)
    sc = _arr(sc, np.int32)
    obc = _arr([1 if b else 0 for b in open_bc], np.int32)
    shells = _arr(shell_d, np.float64)
    n = int(np.prod(sc)) * basis.shape[0]
    col = np.zeros(n, np.int32)
    nc = C.c_int32()
    _check(lib.ether_colour_lattice(_ptr(abc), basis.shape[0], _ptr(basis), _ptr(sc), _ptr(obc), len(shells),
                                    _ptr(shells), float(tol), _ptr(col), C.byref(nc)))
    return col if nc.value > 0 else None


def build_neighbours(abc, basis_xyz, sc, open_bc, shell_d, tol=1e-5, capacity=20):
    """Neighbour lists with the reference's capture rule (see ether_build_neighbours)."""
    lib = load_library()
    abc = _arr(abc, np.float64).reshape(3, 3)
    basis = _arr(basis_xyz, np.float64).reshape(-1, 3)
    sc = _arr(sc, np.int32)
    obc = _arr([1 if b else 0 for b in open_bc], np.int32)
    shells = _arr(shell_d, np.float64)
    n = int(np.prod(sc)) * basis.shape[0]
    cnt = np.zeros(len(shells) * n, np.int32)
    idx = np.full(len(shells) * n * capacity, -1, np.int32)
    rc = lib.ether_build_neighbours(_ptr(abc), basis.shape[0], _ptr(basis), _ptr(sc), _ptr(obc),
                                    len(shells), _ptr(shells), float(tol), int(capacity),
                                    _ptr(cnt), _ptr(idx))
    _check(rc)
    return cnt.reshape(len(shells), n), idx.reshape(len(shells), n, capacity)


def build_spiral_partners(pos, direc, dis, latency_deg, capacity=5):
    """Spiral-state partners of every site (see ether_build_spiral_partners):
    returns (partners (N, capacity) int32, -1 padded; spiral_count)."""
    lib = load_library()
    p = _arr(pos, np.float64).reshape(-1, 3)
    d = _arr(direc, np.float64).reshape(3)
    out = np.full(p.shape[0] * int(capacity), -1, np.int32)
    cnt = C.c_int64()
    rc = lib.ether_build_spiral_partners(p.shape[0], _ptr(p), _ptr(d), float(dis), float(latency_deg),
                                         int(capacity), _ptr(out), C.byref(cnt))
    _check(rc)
    return out.reshape(p.shape[0], int(capacity)), cnt.value


class Engine:
    """One engine context: R replicas of an N-site lattice resident on the GPU."""

    def __init__(self, model: int, n_sites: int, n_replicas: int, seed: int):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.ether_ctx_create(model, n_sites, n_replicas, C.c_uint64(seed), C.byref(h)))
        self.h = h
        self.model = model
        self.N = n_sites
        self.R = n_replicas

    # context management
    def close(self) -> None:
        if self.h:
            self.lib.ether_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc: int) -> None:
        _check(rc, self.h)

    # set-up
    def set_lattice(self, n_species, species, stagger, nn_count, nn_idx, colour=None):
        sp = _arr(species, np.int32)
        st = _arr(stagger, np.float64)
        cnt = _arr(nn_count, np.int32)
        idx = _arr(nn_idx, np.int32)
        n_shells = cnt.shape[0] if cnt.ndim == 2 else 1
        cap = idx.shape[-1]
        col = None if colour is None else _arr(colour, np.int32)
        self._c(self.lib.ether_set_lattice(self.h, int(n_species), _ptr(sp), _ptr(st), n_shells,
                                           cap, _ptr(cnt), _ptr(idx), _ptr(col)))

    def setup(self, n_species, species, stagger, nn_count, nn_idx, j_exc, spin, *,
              zeeman=None, g_factor=2.0, mu_b=0.0, sia=None, angle=False, dphi=0.0,
              compat=0, colour=None, spiral=None):
        """set_lattice + set_exchange + set_zeeman + set_sia + set_proposal
        (+ set_spiral with spiral = partners (N, capacity)) + commit."""
        self.set_lattice(n_species, species, stagger, nn_count, nn_idx, colour)
        j = _arr(j_exc, np.float64)
        s = _arr(spin, np.float64)
        self._c(self.lib.ether_set_exchange(self.h, _ptr(j), _ptr(s)))
        if zeeman is not None:
            hz = _arr(zeeman, np.float64)
            self._c(self.lib.ether_set_zeeman(self.h, 1, _ptr(hz), float(g_factor), float(mu_b)))
        if sia is not None:
            d = _arr(sia, np.float64)
            self._c(self.lib.ether_set_sia(self.h, 1, _ptr(d)))
        self._c(self.lib.ether_set_proposal(self.h, 1 if angle else 0, float(dphi)))
        self._c(self.lib.ether_set_compat(self.h, int(compat)))
        if spiral is not None:
            self.set_spiral(spiral)
        self._c(self.lib.ether_commit(self.h))

    def set_spiral(self, partners) -> None:
        """Spiral-state partners (N, capacity), -1 padded; before commit."""
        pr = _arr(partners, np.int32)
        cap = pr.shape[-1] if pr.size else 0
        self._c(self.lib.ether_set_spiral(self.h, int(cap), _ptr(pr)))
        self.spiral_count = int((pr >= 0).sum())

    def layout(self):
        nc = C.c_int32()
        col = np.zeros(self.N, np.int32)
        rid = np.zeros(self.N, np.int32)
        self._c(self.lib.ether_get_layout(self.h, C.byref(nc), _ptr(col), _ptr(rid)))
        return nc.value, col, rid

    def set_betas(self, beta: Sequence[float]) -> None:
        b = _arr(beta, np.float64)
        assert b.size == self.R
        self._c(self.lib.ether_set_betas(self.h, _ptr(b)))

    def retune_betas(self, beta: Sequence[float]) -> None:
        b = _arr(beta, np.float64)
        assert b.size == self.R
        self._c(self.lib.ether_retune_betas(self.h, _ptr(b)))

    def set_ladders(self, group_start: Sequence[int]) -> None:
        g = _arr(group_start, np.int32)
        self._c(self.lib.ether_set_ladders(self.h, int(g.size - 1), _ptr(g)))

    def set_compat(self, flags: int) -> None:
        self._c(self.lib.ether_set_compat(self.h, int(flags)))

    # hot path
    def init_spins(self, mode: int) -> None:
        self._c(self.lib.ether_init_spins(self.h, int(mode)))

    def sweep(self, n_mcs: int = 1, want_accept: bool = False):
        acc = np.zeros(self.R, np.float64) if want_accept else None
        self._c(self.lib.ether_sweep(self.h, int(n_mcs), _ptr(acc)))
        return acc

    def overrelax(self, passes: float = 1.0) -> None:
        self._c(self.lib.ether_overrelax(self.h, float(passes)))

    def measure(self, sample: bool = False):
        E = np.zeros(self.R, np.float64)
        M = np.zeros(3 * self.R, np.float64)
        self._c(self.lib.ether_measure(self.h, 1 if sample else 0, _ptr(E), _ptr(M)))
        return E, M.reshape(self.R, 3)

    def energy_parts(self):
        parts = np.zeros(4 * self.R, np.float64)
        nc = C.c_int32()
        self._c(self.lib.ether_get_energy_parts(self.h, _ptr(parts), None, C.byref(nc)))
        bonds = None
        if self.model == MODEL_ISING:
            bonds = np.zeros(self.R * max(nc.value, 1), np.int64)
            self._c(self.lib.ether_get_energy_parts(self.h, _ptr(parts), _ptr(bonds), C.byref(nc)))
            bonds = bonds[: self.R * nc.value].reshape(self.R, nc.value)
        return parts.reshape(self.R, 4), bonds

    def moments(self) -> np.ndarray:
        m = np.zeros(self.R * MOM_COUNT, np.float64)
        self._c(self.lib.ether_read_moments(self.h, _ptr(m)))
        return m.reshape(self.R, MOM_COUNT)

    def reset_moments(self) -> None:
        self._c(self.lib.ether_reset_moments(self.h))

    def spiral(self):
        """(sums of psi, psi^2, psi^4 per slot (R, 3), psi of the last measured state (R,))."""
        mom = np.zeros(3 * self.R, np.float64)
        psi = np.zeros(self.R, np.float64)
        self._c(self.lib.ether_read_spiral(self.h, _ptr(mom), _ptr(psi)))
        return mom.reshape(self.R, 3), psi

    def pt_swap(self, parity: int) -> int:
        n = C.c_int32()
        self._c(self.lib.ether_pt_swap(self.h, int(parity), C.byref(n)))
        return n.value

    def pt_state(self):
        slot = np.zeros(self.R, np.int32)
        acc = np.zeros(self.R, np.int64)
        att = np.zeros(self.R, np.int64)
        trips = np.zeros(self.R, np.int64)
        self._c(self.lib.ether_pt_state(self.h, _ptr(slot), _ptr(acc), _ptr(att), _ptr(trips)))
        return slot, acc, att, trips

    def run(self, first_step: int, n_steps: int, mcs_eq: int, coa: int, ovrr_interval: int = 0,
            ovrr_passes: float = 0.0, pt_interval: int = 0, swap_count: int = 0) -> int:
        sc = C.c_int32(swap_count)
        self._c(self.lib.ether_run(self.h, int(first_step), int(n_steps), int(mcs_eq), int(coa),
                                   int(ovrr_interval), float(ovrr_passes), int(pt_interval),
                                   C.byref(sc)))
        return sc.value

    def run_prepare(self, first_step: int, n_steps: int, mcs_eq: int, coa: int, ovrr_interval: int = 0,
                    ovrr_passes: float = 0.0, pt_interval: int = 0) -> None:
        """Capture and upload the graphs run() over the same steps would replay."""
        self._c(self.lib.ether_run_prepare(self.h, int(first_step), int(n_steps), int(mcs_eq), int(coa),
                                           int(ovrr_interval), float(ovrr_passes), int(pt_interval)))

    def set_graphs(self, on: bool) -> None:
        self._c(self.lib.ether_set_graphs(self.h, 1 if on else 0))

    def set_resident(self, on) -> None:
        """Ising sweeps with the lattice resident in LDS (default on where it
        fits): False/0 phase kernels, True/1 chosen, 2 a workgroup per word of
        four replicas, 3 a workgroup per replica."""
        mode = int(on) if not isinstance(on, bool) else (1 if on else 0)
        self._c(self.lib.ether_set_resident(self.h, mode))

    def resident_kind(self) -> int:
        """The Ising sweep path now: 0 phase kernels, 2 a workgroup per word,
        3 a workgroup per replica."""
        k = C.c_int32(0)
        self._c(self.lib.ether_get_resident(self.h, C.byref(k)))
        return int(k.value)

    def set_dft_resident(self, on: bool) -> None:
        """S(q) grid axes of <= 128 cells with the DFT matrices resident in LDS
        (default on); off forces the tiled kernel (bit-identical results)."""
        self._c(self.lib.ether_set_dft_resident(self.h, 1 if on else 0))

    def set_launch(self, sites_per_wave: int, xcd_map: int) -> None:
        self._c(self.lib.ether_set_launch(self.h, int(sites_per_wave), int(xcd_map)))

    def set_lanes(self, lanes_per_site: int) -> None:
        self._c(self.lib.ether_set_lanes(self.h, int(lanes_per_site)))

    AUTOTUNE_SPW = (1, 2, 4, 8, 16)
    AUTOTUNE_MAPS = 3          # xcd_map 0, 1, 2

    def autotune(self, reps: int = 20):
        """Pick the fastest sweep launch shape; returns (sites_per_wave,
        xcd_map, table) with table[xcd_map][k] the ms per sweep of
        sites_per_wave = AUTOTUNE_SPW[k]."""
        spw = C.c_int32()
        xcd = C.c_int32()
        table = np.zeros(self.AUTOTUNE_MAPS * len(self.AUTOTUNE_SPW), np.float64)
        self._c(self.lib.ether_autotune(self.h, int(reps), C.byref(spw), C.byref(xcd), _ptr(table)))
        return spw.value, xcd.value, table.reshape(self.AUTOTUNE_MAPS, -1)

    def set_timing(self, on: bool) -> None:
        self._c(self.lib.ether_set_timing(self.h, 1 if on else 0))

    def timing(self, kind: int):
        """(total ms, launches) of one kernel kind since timing was switched on."""
        ms = C.c_double()
        n = C.c_int64()
        self._c(self.lib.ether_get_timing(self.h, int(kind), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def synchronize(self) -> None:
        self._c(self.lib.ether_synchronize(self.h))

    def get_spins(self, replica: int):
        S = np.zeros(3 * self.N, np.float64)
        phi = np.zeros(self.N, np.float64)
        self._c(self.lib.ether_get_spins(self.h, int(replica), _ptr(S), _ptr(phi)))
        return S.reshape(self.N, 3), phi

    def set_spins(self, replica: int, S, phi=None) -> None:
        s = _arr(S, np.float64).reshape(-1)
        p = None if phi is None else _arr(phi, np.float64)
        self._c(self.lib.ether_set_spins(self.h, int(replica), _ptr(s), _ptr(p)))

    def local_fields(self, replica: int = 0):
        h = np.zeros(3 * self.N, np.float64)
        e = np.zeros(self.N, np.float64)
        self._c(self.lib.ether_local_fields(self.h, int(replica), _ptr(h), _ptr(e)))
        return h.reshape(self.N, 3), e

    def structure_factor(self, sc) -> np.ndarray:
        s = _arr(sc, np.int32)
        out = np.zeros(self.R * self.N, np.float32)
        self._c(self.lib.ether_structure_factor(self.h, _ptr(s), _ptr(out)))
        return out.reshape(self.R, self.N)

    def structure_factor_basis(self, sc, frac) -> np.ndarray:
        """S(q) on the commensurate grid of sc cells with len(frac) basis sites
        at fractional coordinates frac (n_basis, 3): (R, cells)."""
        s = _arr(sc, np.int32)
        f = _arr(frac, np.float64).reshape(-1, 3)
        Q = int(np.prod(s))
        out = np.zeros(self.R * Q, np.float32)
        self._c(self.lib.ether_structure_factor_basis(self.h, _ptr(s), f.shape[0], _ptr(f), _ptr(out)))
        return out.reshape(self.R, Q)

    def structure_factor_line(self, pos, k_vec, q_start, q_stop, dq, noalloc=False) -> np.ndarray:
        """The reference's q-line S(q): (R, n_q, 4) = {sum_a |S_a|^2/N^2, |S_x|/N, |S_y|/N, |S_z|/N};
        noalloc=True calls the allocation-free entry point."""
        p = _arr(pos, np.float64).reshape(-1, 3)
        k = _arr(k_vec, np.float64)
        x = (q_stop - q_start) / dq
        n_q = abs(int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))) + 1
        out = np.zeros(self.R * n_q * 4, np.float32)
        fn = self.lib.ether_structure_factor_line_noalloc if noalloc else self.lib.ether_structure_factor_line
        self._c(fn(self.h, _ptr(p), _ptr(k), float(q_start), float(q_stop), float(dq), int(n_q), _ptr(out)))
        return out.reshape(self.R, n_q, 4)

    # learned ladder
    def mlp_init(self, seed: int) -> None:
        self._c(self.lib.ether_mlp_init(self.h, C.c_uint64(seed)))

    def mlp_params(self) -> np.ndarray:
        p = np.zeros(MLP_NPARAM, np.float32)
        self._c(self.lib.ether_mlp_get_params(self.h, _ptr(p)))
        return p

    def mlp_set_params(self, p) -> None:
        q = _arr(p, np.float32)
        self._c(self.lib.ether_mlp_set_params(self.h, _ptr(q)))

    def mlp_loss_grad(self, x, y, w):
        x = _arr(x, np.float32)
        y = _arr(y, np.float32)
        w = _arr(w, np.float32)
        loss = C.c_float()
        g = np.zeros(MLP_NPARAM, np.float32)
        self._c(self.lib.ether_mlp_loss_grad(self.h, x.size, _ptr(x), _ptr(y), _ptr(w),
                                             C.byref(loss), _ptr(g)))
        return loss.value, g

    def mlp_forward(self, x) -> np.ndarray:
        x = _arr(x, np.float32)
        y = np.zeros(2 * x.size, np.float32)
        self._c(self.lib.ether_mlp_forward(self.h, x.size, _ptr(x), _ptr(y)))
        return y.reshape(x.size, 2)

    def mlp_train(self, x, y, w, iters: int, lr: float) -> float:
        x = _arr(x, np.float32)
        y = _arr(y, np.float32)
        w = _arr(w, np.float32)
        loss = C.c_float()
        self._c(self.lib.ether_mlp_train(self.h, x.size, _ptr(x), _ptr(y), _ptr(w), int(iters),
                                         float(lr), C.byref(loss)))
        return loss.value

    def ladder_place(self, x_mid, x_half, logcv_scale, t_lo, t_hi, n, mode=0):
        """mode: PLACE_LOG_T (input ln T) | PLACE_LENGTH (y1 is the length s(T))."""
        T = np.zeros(int(n), np.float64)
        self._c(self.lib.ether_ladder_place(self.h, float(x_mid), float(x_half), float(logcv_scale),
                                            float(t_lo), float(t_hi), int(n), int(mode), _ptr(T)))
        return T

    def ladder_train(self, temps, kb, first_step, train_mcs, windows, validate_mcs, coa,
                     ovrr_interval, ovrr_passes, pt_interval, learn, placement=0, mlp_seed=1992,
                     swap_count=0):
        """The learned ladder's training and validation windows (see
        ether_ladder_train).  Returns (production ladder, swap count, report[8])."""
        t = _arr(temps, np.float64)
        T = np.zeros(t.size, np.float64)
        rep = np.zeros(8, np.float64)
        sc = C.c_int32(int(swap_count))
        self._c(self.lib.ether_ladder_train(self.h, t.size, _ptr(t), float(kb), int(first_step),
                                            int(train_mcs), int(windows), int(validate_mcs), int(coa),
                                            int(ovrr_interval), float(ovrr_passes), int(pt_interval),
                                            1 if learn else 0, int(placement), C.c_uint64(int(mlp_seed)),
                                            C.byref(sc), _ptr(T), _ptr(rep)))
        return T, sc.value, rep


class Batch:
    """Several engines advanced together by one launch per kernel kind
    (ether_batch_*): each engine is one lattice with its own replicas,
    temperatures and ladders; results are those of Engine.run on each."""

    def __init__(self, engines):
        self.lib = load_library()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        h = C.c_void_p()
        _check(self.lib.ether_batch_create(C.cast(arr, C.c_void_p), len(self.engines), C.byref(h)))
        self.h = h

    def run(self, first_step: int, n_steps: int, mcs_eq: int, coa: int, ovrr_interval: int = 0,
            ovrr_passes: float = 0.0, pt_interval: int = 0, swap_count: int = 0) -> int:
        sc = C.c_int32(swap_count)
        _check(self.lib.ether_batch_run(self.h, int(first_step), int(n_steps), int(mcs_eq), int(coa),
                                        int(ovrr_interval), float(ovrr_passes), int(pt_interval),
                                        C.byref(sc)))
        return sc.value

    def set_graphs(self, on: bool) -> None:
        _check(self.lib.ether_batch_set_graphs(self.h, 1 if on else 0))

    def resident(self) -> bool:
        """Whether the last run swept with the spins resident in LDS."""
        on = C.c_int32(0)
        _check(self.lib.ether_batch_get_resident(self.h, C.byref(on)))
        return bool(on.value)

    def close(self) -> None:
        if self.h:
            self.lib.ether_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
