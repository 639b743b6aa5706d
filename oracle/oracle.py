"""ctypes binding of the CPU FP64 oracle (oracle/libether_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product path (ether_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libether_oracle.so")
_lib = None

_vp = C.c_void_p
NPARAM = 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2
RNG_METROPOLIS, RNG_INIT, RNG_OVRR, RNG_PT, RNG_START = 1, 2, 3, 4, 5


class OModel(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("n_sites", C.c_int32), ("n_species", C.c_int32),
        ("n_shells", C.c_int32), ("capacity", C.c_int32),
        ("species", _vp), ("stagger", _vp), ("nn_count", _vp), ("nn_idx", _vp),
        ("j_exc", _vp), ("spin", _vp),
        ("zeeman", C.c_int32), ("h", C.c_double * 3), ("g_factor", C.c_double), ("mu_b", C.c_double),
        ("sia", C.c_int32), ("sia_vec", _vp),
        ("angle", C.c_int32), ("dphi", C.c_double),
    ]


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        L.o_philox4x32_10.argtypes = [_vp, _vp, _vp]
        L.o_rng.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _vp]
        L.o_neighbours_brute.restype = C.c_int
        L.o_neighbours_brute.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, _vp, C.c_double,
                                         C.c_int32, _vp, _vp]
        for n in ("o_site_energy",):
            getattr(L, n).restype = C.c_double
            getattr(L, n).argtypes = [C.POINTER(OModel), _vp, C.c_int32]
        L.o_delta_energy.restype = C.c_double
        L.o_delta_energy.argtypes = [C.POINTER(OModel), _vp, C.c_int32, _vp]
        L.o_exchange_field.argtypes = [C.POINTER(OModel), _vp, C.c_int32, _vp]
        L.o_total_energy.argtypes = [C.POINTER(OModel), _vp, _vp]
        L.o_ising_bond_sums.argtypes = [C.POINTER(OModel), _vp, _vp, C.c_int32, _vp]
        L.o_magnetisation.argtypes = [C.POINTER(OModel), _vp, _vp]
        L.o_propose.argtypes = [C.POINTER(OModel), C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_uint32, _vp, C.c_double, _vp, _vp, _vp]
        L.o_propose_site.argtypes = [C.POINTER(OModel), C.c_uint64, C.c_uint32, C.c_int32,
                                     C.c_uint32, _vp, _vp, _vp, _vp, _vp]
        L.o_sweep_batch.restype = C.c_int64
        L.o_sweep_batch.argtypes = [C.POINTER(OModel), _vp, _vp, C.c_int32, _vp, C.c_int32, _vp,
                                    _vp, C.c_uint64, C.c_uint32, _vp]
        L.o_start_state.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp]
        L.o_ising_threshold.restype = C.c_uint32
        L.o_ising_threshold.argtypes = [C.c_double, C.c_double]
        L.o_sweep.restype = C.c_int64
        L.o_sweep.argtypes = [C.POINTER(OModel), _vp, _vp, _vp, C.c_int32, _vp, C.c_double,
                              C.c_uint64, C.c_uint32, C.c_uint32]
        L.o_reflect.argtypes = [C.POINTER(OModel), _vp, C.c_int32]
        L.o_overrelax_pass.argtypes = [C.POINTER(OModel), _vp, _vp, _vp, C.c_int32, _vp, C.c_double,
                                       C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        L.o_run_chain.argtypes = [C.POINTER(OModel), _vp, _vp, _vp, C.c_int32, _vp, C.c_double,
                                  C.c_uint64, C.c_uint32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_double, _vp]
        L.o_repeat_observables.argtypes = [_vp, C.c_double, C.c_int32, _vp]
        L.o_over_repeats.argtypes = [_vp, C.c_int32, _vp, _vp]
        L.o_pt_accept.restype = C.c_int
        L.o_spiral_partners.restype = C.c_int64
        L.o_spiral_partners.argtypes = [C.c_int32, _vp, _vp, C.c_double, C.c_double, C.c_int32, _vp]
        L.o_spiral_psi.restype = C.c_double
        L.o_spiral_psi.argtypes = [C.c_int32, _vp, _vp, C.c_int32, C.c_int64]
        L.o_spiral_observables.argtypes = [_vp, C.c_double, C.c_double, C.c_int64, _vp]
        L.o_pt_accept.argtypes = [C.c_double] * 5
        L.o_mlp_forward.argtypes = [_vp, C.c_int32, _vp, _vp]
        L.o_mlp_loss_grad.restype = C.c_double
        L.o_mlp_loss_grad.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp]
        L.o_structure_factor.restype = C.c_double
        L.o_structure_factor.argtypes = [C.c_int32, _vp, _vp, _vp]
        L.o_structure_factor_parts.argtypes = [C.c_int32, _vp, _vp, _vp, _vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _a(x, dt):
    return np.ascontiguousarray(np.asarray(x, dtype=dt))


def philox(ctr, key):
    c = _a(ctr, np.uint32)
    k = _a(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().o_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def rng(seed, id_, replica, step, purpose, attempt=0):
    out = np.zeros(4, np.uint32)
    lib().o_rng(C.c_uint64(seed), id_, replica, step, purpose, attempt, _p(out))
    return out


def start_state(seed, model, angle, file_index):
    """generate_supercell start state of the basis sites: (S [nb, 3], phi [nb])."""
    fi = _a(file_index, np.int32)
    S = np.zeros((fi.size, 3))
    phi = np.zeros(fi.size)
    lib().o_start_state(C.c_uint64(seed), int(model), int(angle), fi.size, _p(fi), _p(S), _p(phi))
    return S, phi


def neighbours_brute(abc, basis_xyz, sc, open_bc, shells, tol=1e-5, capacity=20):
    abc = _a(abc, np.float64)
    basis = _a(basis_xyz, np.float64).reshape(-1, 3)
    sc = _a(sc, np.int32)
    obc = _a([1 if b else 0 for b in open_bc], np.int32)
    sh = _a(shells, np.float64)
    n = int(np.prod(sc)) * basis.shape[0]
    cnt = np.zeros(len(sh) * n, np.int32)
    idx = np.full(len(sh) * n * capacity, -1, np.int32)
    rc = lib().o_neighbours_brute(_p(abc), basis.shape[0], _p(basis), _p(sc), _p(obc), len(sh),
                                  _p(sh), tol, capacity, _p(cnt), _p(idx))
    if rc != 0:
        raise ValueError(f"o_neighbours_brute failed ({rc})")
    return cnt.reshape(len(sh), n), idx.reshape(len(sh), n, capacity)


class Oracle:
    """Holds an o_model and the arrays it points to."""

    def __init__(self, model, n_sites, n_species, species, stagger, nn_count, nn_idx, j_exc, spin,
                 zeeman=None, g_factor=2.0, mu_b=0.0, sia=None, angle=False, dphi=0.0):
        self.keep = dict(
            species=_a(species, np.int32), stagger=_a(stagger, np.float64),
            nn_count=_a(nn_count, np.int32), nn_idx=_a(nn_idx, np.int32),
            j_exc=_a(j_exc, np.float64), spin=_a(spin, np.float64),
            sia=_a(sia, np.float64) if sia is not None else np.zeros(3 * n_species))
        k = self.keep
        cnt = k["nn_count"]
        m = OModel()
        m.model = model
        m.n_sites = n_sites
        m.n_species = n_species
        m.n_shells = cnt.size // n_sites
        m.capacity = k["nn_idx"].size // cnt.size
        m.species, m.stagger = _p(k["species"]), _p(k["stagger"])
        m.nn_count, m.nn_idx = _p(cnt), _p(k["nn_idx"])
        m.j_exc, m.spin = _p(k["j_exc"]), _p(k["spin"])
        m.zeeman = 1 if zeeman is not None else 0
        hh = zeeman if zeeman is not None else (0.0, 0.0, 0.0)
        for x in range(3):
            m.h[x] = float(hh[x])
        m.g_factor, m.mu_b = float(g_factor), float(mu_b)
        m.sia = 1 if sia is not None else 0
        m.sia_vec = _p(k["sia"])
        m.angle = 1 if angle else 0
        m.dphi = float(dphi)
        self.m = m
        self.N = n_sites

    @classmethod
    def from_model(cls, model):
        t = model.tags
        return cls(model.model_id, model.n_sites, model.n_species, model.species, model.stagger,
                   model.nn_count, model.nn_idx, model.j_exc.reshape(-1), model.spin,
                   zeeman=model.h if t.zeeman else None, g_factor=model.g_factor, mu_b=model.mu_b,
                   sia=model.sia_vec.reshape(-1) if model.sia_vec is not None else None,
                   angle=t.angle, dphi=t.dphi)

    def site_energy(self, S, i):
        S = _a(S, np.float64).reshape(-1)
        return lib().o_site_energy(C.byref(self.m), _p(S), int(i))

    def delta_energy(self, S, i, s_new):
        S = _a(S, np.float64).reshape(-1)
        sn = _a(s_new, np.float64)
        return lib().o_delta_energy(C.byref(self.m), _p(S), int(i), _p(sn))

    def exchange_field(self, S, i):
        S = _a(S, np.float64).reshape(-1)
        h = np.zeros(3)
        lib().o_exchange_field(C.byref(self.m), _p(S), int(i), _p(h))
        return h

    def total_energy(self, S):
        S = _a(S, np.float64).reshape(-1)
        parts = np.zeros(3)
        lib().o_total_energy(C.byref(self.m), _p(S), _p(parts))
        return parts

    def ising_bond_sums(self, S, class_of_edge, n_classes):
        S = _a(S, np.float64).reshape(-1)
        c = _a(class_of_edge, np.int32)
        b = np.zeros(n_classes, np.int64)
        lib().o_ising_bond_sums(C.byref(self.m), _p(S), _p(c), n_classes, _p(b))
        return b

    def magnetisation(self, S):
        S = _a(S, np.float64).reshape(-1)
        M = np.zeros(3)
        lib().o_magnetisation(C.byref(self.m), _p(S), _p(M))
        return M

    def propose(self, seed, id_, replica, step, s_old, phi_old=0.0, purpose=RNG_METROPOLIS):
        so = _a(s_old, np.float64)
        sn = np.zeros(3)
        ph = C.c_double()
        eta = C.c_double()
        lib().o_propose(C.byref(self.m), C.c_uint64(seed), id_, replica, step, purpose, _p(so),
                        float(phi_old), _p(sn), C.byref(ph), C.byref(eta))
        return sn, ph.value, eta.value

    def propose_site(self, seed, id_, step, s_old, phi_old=None):
        """Proposals of one site for all R replicas (device rule incl. pooling)."""
        so = _a(s_old, np.float64).reshape(-1, 3)
        R = so.shape[0]
        po = None if phi_old is None else _a(phi_old, np.float64)
        sn = np.zeros((R, 3))
        pn = np.zeros(R)
        eta = np.zeros(R)
        lib().o_propose_site(C.byref(self.m), C.c_uint64(seed), id_, R, step, _p(so), _p(po),
                             _p(sn), _p(pn), _p(eta))
        return sn, pn, eta

    def sweep_batch(self, S, phi, order, rng_id, betas, seed, step):
        """In place on S (R,N,3) float64 and phi (R,N) (or None); returns accepted per replica."""
        assert S.dtype == np.float64 and S.flags.c_contiguous
        R = S.shape[0]
        o = _a(order, np.int32)
        r = _a(rng_id, np.int32)
        b = _a(betas, np.float64)
        acc = np.zeros(R, np.int64)
        lib().o_sweep_batch(C.byref(self.m), _p(S), _p(phi), R, _p(o), o.size, _p(r), _p(b),
                            C.c_uint64(seed), step, _p(acc))
        return acc

    def sweep(self, S, phi, order, rng_id, beta, seed, replica, step):
        """In place on S (N,3) float64 and phi (N,) float64 (or None)."""
        assert S.dtype == np.float64 and S.flags.c_contiguous
        o = _a(order, np.int32)
        r = _a(rng_id, np.int32)
        return lib().o_sweep(C.byref(self.m), _p(S), _p(phi), _p(o), o.size, _p(r), float(beta),
                             C.c_uint64(seed), replica, step)

    def run_chain(self, S, phi, order, rng_id, beta, seed, replica, mcs, eq, coa, ovrr_iv=0,
                  ovrr_p=0.0):
        """One chain through the reference's MCS loop; returns the 8-moment record."""
        o = _a(order, np.int32)
        r = _a(rng_id, np.int32)
        mom = np.zeros(8)
        lib().o_run_chain(C.byref(self.m), _p(S), _p(phi), _p(o), o.size, _p(r), float(beta),
                          C.c_uint64(seed), replica, mcs, eq, coa, ovrr_iv, float(ovrr_p), _p(mom))
        return mom

    def reflect(self, S, i):
        lib().o_reflect(C.byref(self.m), _p(S), int(i))

    def overrelax_pass(self, S, phi, order, rng_id, frac, seed, replica, step, pass_):
        o = _a(order, np.int32)
        r = _a(rng_id, np.int32)
        lib().o_overrelax_pass(C.byref(self.m), _p(S), _p(phi), _p(o), o.size, _p(r), float(frac),
                               C.c_uint64(seed), replica, step, pass_)


def ising_threshold(beta, dE):
    return lib().o_ising_threshold(float(beta), float(dE))


def repeat_observables(mom7, beta, n_sites):
    m = _a(mom7, np.float64)
    out = np.zeros(6)
    lib().o_repeat_observables(_p(m), float(beta), int(n_sites), _p(out))
    return out


def over_repeats(per):
    p = _a(per, np.float64)
    mean = np.zeros(6)
    sem = np.zeros(6)
    lib().o_over_repeats(_p(p), p.shape[0], _p(mean), _p(sem))
    return mean, sem


def spiral_partners(pos, direc, dis, latency_deg, cap=5):
    """(partners (N, cap) int32 -1 padded, spiral_count); count -1 on overflow."""
    p = _a(pos, np.float64).reshape(-1, 3)
    d = _a(direc, np.float64).reshape(3)
    out = np.full(p.shape[0] * cap, -1, np.int32)
    cnt = lib().o_spiral_partners(p.shape[0], _p(p), _p(d), float(dis), float(latency_deg), int(cap),
                                  _p(out))
    return out.reshape(p.shape[0], cap), int(cnt)


def spiral_psi(S, partners, count):
    s = _a(S, np.float64).reshape(-1, 3)
    pr = _a(partners, np.int32)
    return lib().o_spiral_psi(s.shape[0], _p(s), _p(pr), pr.shape[-1], int(count))


def spiral_observables(sums3, n_samples, beta, count):
    s = _a(sums3, np.float64)
    out = np.zeros(3)
    lib().o_spiral_observables(_p(s), float(n_samples), float(beta), int(count), _p(out))
    return out


def pt_accept(beta, beta_p, E, E_p, u):
    return bool(lib().o_pt_accept(beta, beta_p, E, E_p, u))


def mlp_forward(params, x):
    p = _a(params, np.float64)
    x = _a(x, np.float64)
    y = np.zeros(2 * x.size)
    lib().o_mlp_forward(_p(p), x.size, _p(x), _p(y))
    return y.reshape(-1, 2)


def mlp_loss_grad(params, x, y, w):
    p = _a(params, np.float64)
    x = _a(x, np.float64)
    y = _a(y, np.float64)
    w = _a(w, np.float64)
    g = np.zeros(NPARAM)
    loss = lib().o_mlp_loss_grad(_p(p), x.size, _p(x), _p(y), _p(w), _p(g))
    return loss, g


def structure_factor(pos, S, q):
    pos = _a(pos, np.float64)
    S = _a(S, np.float64)
    q = _a(q, np.float64)
    return lib().o_structure_factor(pos.shape[0], _p(pos), _p(S), _p(q))


def structure_factor_parts(pos, S, q):
    """{sum_a |S_a(q)|^2/N^2, |S_x|/N, |S_y|/N, |S_z|/N} by the FP64 direct sum."""
    p = _a(pos, np.float64).reshape(-1, 3)
    s = _a(S, np.float64).reshape(-1, 3)
    qq = _a(q, np.float64)
    out = np.zeros(4)
    lib().o_structure_factor_parts(p.shape[0], _p(p), _p(s), _p(qq), _p(out))
    return out
