"""The reference's result files, written from a finished batched run.

ETHER keeps, per temperature, the observables over repeats and the spin
configuration its rank holds at the end of the last repeat; it stores them in
two direct-access files and derives the text outputs from those
(src/process_observables.f90:34-127, src/ether.f90:216-231,
src/write_output_files.f90, src/write_gss.f90, src/write_spins_at_K.f90):

  ETHER.obs                 record t: 20 + 3*nspecies FP64 observables
  ETHER.spn                 record t: ion(0:8, 1:N) FP64, site by site
                            (id, Sx, Sy, Sz, species, phi, x, y, z)
  gss.dat                   lattice header, then per t the spins and species
  moment_vectors.dat        per t the staggered moment per included species
  <T>spK.xsf                per t the spins as an XCrySDen crystal
  energy.dat, magnetization.dat, spiral_state_parameters.dat  (driver.Results)

Direct-access unformatted records carry no markers: record t (1-based) is
the t-th block of recl bytes, little-endian doubles in Fortran column order.

In the batched run replica r = k*nscan + t is repeat k of temperature t.  The
configuration "at temperature t after the last repeat" is the replica whose
slot is (repeat-1)*nscan + t: with tempering on, labels move between slots,
as the reference's spins move between ranks.
"""
from __future__ import annotations

import os
from typing import Sequence

import numpy as np

from .inputs import Model

TOTAL_OBSERVABLES = 20          # src/ether.f90:105


def final_states(eng, nscan: int, repeat: int):
    """Spins and phi at each temperature after the last repeat: (nscan, N, 3), (nscan, N)."""
    slot_of = np.asarray(eng.pt_state()[0])
    rep_at = np.empty_like(slot_of)
    rep_at[slot_of] = np.arange(slot_of.size)
    S = np.zeros((nscan, eng.N, 3))
    phi = np.zeros((nscan, eng.N))
    for t in range(nscan):
        S[t], phi[t] = eng.get_spins(int(rep_at[(repeat - 1) * nscan + t]))
    return S, phi


def moment_vectors(model: Model, S: np.ndarray) -> np.ndarray:
    """get_moment_vectors (src/get_moment_vectors.f90:20-37): per species
    sum of stagger_i * s(species) * S_i over the sites; (n_species, 3)."""
    mm = np.zeros((model.n_species, 3))
    v = (model.stagger[:, None] * S) * model.spin[model.species - 1][:, None]
    np.add.at(mm, model.species - 1, v)     # site order, as the reference's loop
    return mm


def species_totals(model: Model) -> np.ndarray:
    """tions(1:nspecies): sites per species in the supercell (src/generate_supercell.f90:142-147)."""
    return np.bincount(model.species, minlength=model.n_species + 1)[1:]


def obs_table(res, model: Model, mm: np.ndarray) -> np.ndarray:
    """Records of ETHER.obs in the order of process_observables('store')
    (src/process_observables.f90:40-75), one row per temperature."""
    nscan = len(res.temps)
    tab = np.zeros((nscan, TOTAL_OBSERVABLES + 3 * model.n_species))
    for t in range(nscan):
        m, s = res.mean[t], res.sem[t]
        row = [res.temps[t], m[3], m[5], s[3], s[5], m[4], s[4],
               m[0], m[2], s[0], s[2], m[1], s[1], res.accept_pct[t]]
        if res.spiral_mean is not None:
            p, q = res.spiral_mean[t], res.spiral_sem[t]
            row += [p[0], p[2], q[0], q[2], p[1], q[1]]
        else:
            row += [0.0] * 6
        tab[t, :TOTAL_OBSERVABLES] = row
        tab[t, TOTAL_OBSERVABLES:] = mm[t].reshape(-1)
    return tab


def spn_records(model: Model, S: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """ion(0:8, 1:N) per temperature (src/ether.f90:228-231): (nscan, N, 9)."""
    nscan, N = S.shape[0], model.n_sites
    rec = np.zeros((nscan, N, 9))
    rec[:, :, 0] = np.arange(1, N + 1)
    rec[:, :, 1:4] = S
    rec[:, :, 4] = model.species
    rec[:, :, 5] = phi
    rec[:, :, 6:9] = model.positions
    return rec


def write_direct(path: str, records: np.ndarray) -> None:
    """A direct-access unformatted file of equal FP64 records (src/rw_file.f90:35-55)."""
    np.ascontiguousarray(records, dtype="<f8").tofile(path)


def read_direct(path: str, record_doubles: int) -> np.ndarray:
    return np.fromfile(path, dtype="<f8").reshape(-1, record_doubles)


def _f(x: float, w: int, d: int) -> str:
    return f"{x:{w}.{d}f}"


def gss_text(model: Model, temps: Sequence[float], S: np.ndarray) -> str:
    """gss.dat (src/write_gss.f90:34-78)."""
    t = model.tags
    sc = model.sc
    out = [f"{len(temps):5d} {sc[0]:5d} {sc[1]:5d} {sc[2]:5d}",
           f"{1:5d} {1:5d} {1:5d} {sc[0]:5d} {sc[1]:5d} {sc[2]:5d}",
           f"{len(model.basis_species):5d} {model.n_species:5d}{_f(t.ht, 11, 5)} {_f(t.lt, 11, 5)} "
           f"{_f(t.tint, 11, 5)}",
           (" KbT/J" if t.para else " K").ljust(20)]
    tions = species_totals(model)
    for k in range(0, len(tions), 5):
        out.append(" ".join(f"{n:8d}" for n in tions[k:k + 5]))
    for i in range(3):
        out.append(" ".join(_f(v, 10, 5) for v in model.structure.abc[i]))
    out.append(" # lattice points")
    p = model.positions
    for i in range(model.n_sites):
        out.append(f"{_f(p[i, 0], 11, 6)} {_f(p[i, 1], 11, 6)} {_f(p[i, 2], 11, 6)} {i + 1:8d}")
    for k, T in enumerate(temps):
        out.append(_f(T, 11, 5))
        s = S[k]
        for i in range(model.n_sites):
            out.append(f"{_f(s[i, 0], 11, 6)} {_f(s[i, 1], 11, 6)} {_f(s[i, 2], 11, 6)} "
                       f"{int(model.species[i]):8d}")
    return "\n".join(out) + "\n"


def included_species(model: Model):
    """included_species_ID in the order of the `species` tag (src/generate_supercell.f90:65-74)."""
    names = list(model.structure.species)
    return [names.index(n) + 1 for n in model.tags.species if n in names]


def moment_vectors_text(model: Model, temps: Sequence[float], mm: np.ndarray) -> str:
    """moment_vectors.dat (src/write_output_files.f90:66-80, format 101 (f11.5,2x,500f9.3))."""
    inc = included_species(model)
    names = list(model.structure.species)
    out = [" # Showing moment vectors (Mx, My, Mz) for species(sp):"]
    for k in range(0, len(inc), 5):
        out.append(" #\t" + "".join(names[j - 1].ljust(2)[:2].rjust(3) for j in inc[k:k + 5]))
    out.append(" # Temp, sp1(Mx,My,Mz), sp1(Mx,My,Mz),..spN(Mx,My,Mz), sp1|M|, sp2|M|,..spN|M|")
    tions = species_totals(model)
    for k, T in enumerate(temps):
        vals = []
        for j in inc:
            vals += list(mm[k, j - 1] / tions[j - 1])
        for j in inc:
            vals.append(np.linalg.norm(mm[k, j - 1]) / tions[j - 1])
        out.append(_f(T, 11, 5) + "  " + "".join(_f(v, 9, 3) for v in vals))
    return "\n".join(out) + "\n"


def spk_name(T: float) -> str:
    """write(filename, '(f9.4,"spK.xsf")') K, then adjustl (src/write_spins_at_K.f90:36-39)."""
    return (_f(T, 9, 4) + "spK.xsf").strip()


def spk_text(model: Model, S: np.ndarray) -> str:
    """<T>spK.xsf (src/write_spins_at_K.f90:33-66).  The header lines are
    list-directed in the reference; here each value is written in full
    precision, which XCrySDen reads the same way."""
    abc = np.asarray(model.structure.abc, np.float64)
    sc = model.sc

    def vec(v):
        return " " + " ".join(f"{x:.16g}" for x in v)

    prim = [vec(abc[a] * sc[a]) for a in range(3)]
    out = [" CRYSTAL", " PRIMVEC", *prim, " CONVEC", *prim, " PRIMCOORD", f" {model.n_sites} 1"]
    names = list(model.structure.species)
    p = model.positions
    for i in range(model.n_sites):
        name = names[model.species[i] - 1].ljust(2)[:2].rjust(5)
        out.append(name + " " + " ".join(_f(v, 13, 7) for v in (*p[i], *S[i])))
    return "\n".join(out) + "\n"


def write_all(outdir: str, res, model: Model) -> None:
    """ETHER.obs, ETHER.spn, gss.dat, moment_vectors.dat and the spK files."""
    S, phi = res.final_S, res.final_phi
    mm = np.stack([moment_vectors(model, S[t]) for t in range(S.shape[0])])
    write_direct(os.path.join(outdir, "ETHER.obs"), obs_table(res, model, mm))
    write_direct(os.path.join(outdir, "ETHER.spn"), spn_records(model, S, phi))
    with open(os.path.join(outdir, "gss.dat"), "w") as f:
        f.write(gss_text(model, res.temps, S))
    with open(os.path.join(outdir, "moment_vectors.dat"), "w") as f:
        f.write(moment_vectors_text(model, res.temps, mm))
    for t, T in enumerate(res.temps):
        with open(os.path.join(outdir, spk_name(T)), "w") as f:
            f.write(spk_text(model, S[t]))
