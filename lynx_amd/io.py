"""The two files the reference's SCF driver writes and reads back.

Density restart file (src/io/DensityIO.hpp:11-43, DensityIO.cpp:7-121):
a 128-byte little-endian header

    offset  0  char[8]   magic "LYNX_RHO"
            8  uint32    version (1)
           12  uint32    nspin
           16  uint32    nx, ny, nz
           28  uint32    padding (0)
           32  double[9] lattice vectors, row i = vector i (Bohr)
          104  double    dV (Bohr^3)
          112  16 zero bytes

followed by nspin blocks of Nx Ny Nz doubles, x fastest (the grid's own
order), one block per spin: the total density for one spin, spin up then
spin down for two.  Reading checks the magic, the version, the grid size
and the spin count against the calculation and gives up (None, the
reference's false) on any mismatch, as DensityIO::read does; the lattice
and dV in the header are informational.

Result JSON (src/io/ResultWriter.cpp:36-205, ResultWriter.hpp:14-137): the
sections system, cell, grid, electronic, kpoints, scf, energies_Ha,
forces_Ha_per_bohr, stress, eigenvalues and timing_s with the reference's
key names and defaults, written with two-space indentation and keys in
sorted order (the reference's JSON objects are ordered maps).
``ResultWriter.from_calculator`` fills it from a finished
``lynx_amd.scf.Calculator`` run the way the reference's end-to-end driver
does (tests/e2e/test_EndToEnd.cpp:312-470).

Plain host-side file handling: nothing here touches the GPU.
"""
from __future__ import annotations

import copy
import json
import math
import os
import struct
import time

import numpy as np

DENSITY_MAGIC = b"LYNX_RHO"
DENSITY_VERSION = 1
_HEADER = struct.Struct("<8s6I9dd16x")
assert _HEADER.size == 128

AU_TO_GPA = 29421.01569650548   # Ha/Bohr^3 in GPa (test_EndToEnd.cpp:444)


def write_density(path, rho, shape, latvec, dV):
    """Write ``rho`` (nspin x Nd, or Nd for one spin; x fastest) on the
    grid ``shape`` = (Nx, Ny, Nz) of the cell ``latvec`` (rows are the
    lattice vectors) with volume element ``dV``.  Returns False when the
    file cannot be written (DensityIO::write, DensityIO.cpp:7-49)."""
    nx, ny, nz = (int(v) for v in shape)
    nd = nx * ny * nz
    r = np.ascontiguousarray(rho, dtype="<f8")
    if r.ndim == 1:
        r = r.reshape(1, -1)
    if r.ndim != 2 or r.shape[1] != nd or r.shape[0] not in (1, 2):
        raise ValueError(f"write_density: rho of shape {np.shape(rho)} is not nspin x {nd}")
    lv = np.asarray(latvec, float).reshape(3, 3)
    hdr = _HEADER.pack(DENSITY_MAGIC, DENSITY_VERSION, r.shape[0], nx, ny, nz, 0, *lv.ravel(), float(dV))
    try:
        with open(path, "wb") as f:
            f.write(hdr)
            f.write(r.tobytes())
    except OSError:
        return False
    return True


def density_header(path):
    """The header fields of a density file as a dict (no validation)."""
    with open(path, "rb") as f:
        raw = f.read(_HEADER.size)
    if len(raw) != _HEADER.size:
        raise ValueError(f"{path}: shorter than the 128-byte header")
    v = _HEADER.unpack(raw)
    return {"magic": v[0], "version": v[1], "nspin": v[2], "shape": (v[3], v[4], v[5]),
            "latvec": np.array(v[7:16]).reshape(3, 3), "dV": v[16]}


def read_density(path, shape, nspin, magic=DENSITY_MAGIC):
    """The densities (nspin x Nd) of a density file written for the grid
    ``shape`` and ``nspin`` spins, or None when the file is missing, short,
    or its magic, version, grid or spin count differ (DensityIO::read,
    DensityIO.cpp:51-121).  ``magic`` names the accepted tag; files from
    earlier builds of the reference carry "SPARCRHO" with the same layout."""
    try:
        with open(path, "rb") as f:
            raw = f.read(_HEADER.size)
            if len(raw) != _HEADER.size:
                return None
            v = _HEADER.unpack(raw)
            if v[0] != magic or v[1] != DENSITY_VERSION:
                return None
            if tuple(v[3:6]) != tuple(int(s) for s in shape) or v[2] != int(nspin):
                return None
            nd = v[3] * v[4] * v[5]
            body = f.read(8 * nd * v[2])
    except OSError:
        return None
    if len(body) != 8 * nd * v[2]:
        return None
    return np.frombuffer(body, dtype="<f8").astype(float).reshape(v[2], nd)


def total_and_mag(rho):
    """Total density and magnetization of nspin x Nd spin densities, as
    DensityIO::read recomputes them (DensityIO.cpp:104-118): for one spin
    the total is the block itself and the magnetization None."""
    r = np.asarray(rho, float)
    if r.shape[0] == 1:
        return r[0].copy(), None
    return r[0] + r[1], r[0] - r[1]


# ------------------------------------------------------------------ results

def _defaults():
    # ResultWriter.hpp:14-137
    return {
        "system": {"n_atoms": 0, "n_electrons": 0, "atom_types": [], "atoms_cartesian_bohr": [],
                   "atoms_fractional": []},
        "cell": {"lattice_vectors_bohr": [[0.0] * 3 for _ in range(3)], "volume_bohr3": 0.0, "orthogonal": True},
        "grid": {"Nx": 0, "Ny": 0, "Nz": 0, "dx_bohr": 0.0, "dy_bohr": 0.0, "dz_bohr": 0.0, "fd_order": 0,
                 "boundary_conditions": ["P", "P", "P"]},
        "electronic": {"xc": "", "spin": "none", "soc": False, "smearing": "", "electronic_temperature_K": 0.0,
                       "kBT_Ha": 0.0, "n_states": 0, "cheb_degree": 0},
        "kpoints": {"grid": [1, 1, 1], "shift": [0.0, 0.0, 0.0], "list": []},
        "scf": {"max_iter": 0, "tolerance": 0.0, "mixing_variable": "", "preconditioner": "", "mixing_history": 0,
                "mixing_parameter": 0.0, "converged": False, "n_iterations": 0, "history": []},
        "energies_Ha": {"total_free_energy": 0.0, "free_energy_per_atom": 0.0, "band_structure": 0.0,
                        "exchange_correlation": 0.0, "hartree": 0.0, "self_energy": 0.0, "correction_energy": 0.0,
                        "self_plus_correction": 0.0, "entropy_term": 0.0, "fermi_level": 0.0},
        "forces_Ha_per_bohr": {"max": 0.0, "rms": 0.0, "per_atom": []},
        "stress": {"computed": False, "tensor_Ha_per_bohr3": [[0.0] * 3 for _ in range(3)],
                   "tensor_GPa": [[0.0] * 3 for _ in range(3)], "voigt_Ha_per_bohr3": [0.0] * 6,
                   "voigt_GPa": [0.0] * 6, "pressure_GPa": 0.0, "max_abs_stress_GPa": 0.0},
        "eigenvalues": {"units": "Ha", "n_states": 0, "per_kpoint": []},
        "timing_s": {"scf_total": 0.0, "force": 0.0, "stress": 0.0, "walltime_total": 0.0},
    }


def _plain(x):
    # numpy scalars and arrays to JSON values; non-finite numbers to null
    # (what the reference's JSON library writes for them)
    if isinstance(x, dict):
        return {k: _plain(v) for k, v in x.items()}
    if isinstance(x, (list, tuple, np.ndarray)):
        return [_plain(v) for v in (x.tolist() if isinstance(x, np.ndarray) else x)]
    if isinstance(x, (bool, np.bool_)):
        return bool(x)
    if isinstance(x, (int, np.integer)):
        return int(x)
    if isinstance(x, (float, np.floating)):
        x = float(x)
        return x if math.isfinite(x) else None
    return x


class ResultWriter:
    """The reference's result JSON.  Each ``set_<section>(**fields)``
    updates one section (unknown keys are an error); ``write(path)`` adds
    ``timestamp_utc`` and writes the file, creating its directory."""

    _SECTIONS = ("system", "cell", "grid", "electronic", "kpoints", "scf", "energies_Ha", "forces_Ha_per_bohr",
                 "stress", "eigenvalues", "timing_s")

    def __init__(self):
        self.lynx_version = ""
        self.test_name = ""
        self.input_file = ""
        self.device = ""
        self.mpi_ranks = 1
        self.sections = _defaults()

    def _set(self, name, fields):
        sec = self.sections[name]
        for k, v in fields.items():
            if k not in sec:
                raise KeyError(f"ResultWriter.{name}: unknown field {k!r}")
            sec[k] = v
        return self

    def set_system(self, **kw): return self._set("system", kw)
    def set_cell(self, **kw): return self._set("cell", kw)
    def set_grid(self, **kw): return self._set("grid", kw)
    def set_electronic(self, **kw): return self._set("electronic", kw)
    def set_kpoints(self, **kw): return self._set("kpoints", kw)
    def set_scf(self, **kw): return self._set("scf", kw)
    def set_energies(self, **kw): return self._set("energies_Ha", kw)
    def set_forces(self, **kw): return self._set("forces_Ha_per_bohr", kw)
    def set_stress(self, **kw): return self._set("stress", kw)
    def set_eigenvalues(self, **kw): return self._set("eigenvalues", kw)
    def set_timing(self, **kw): return self._set("timing_s", kw)

    def as_dict(self):
        d = {"lynx_version": self.lynx_version, "test_name": self.test_name, "input_file": self.input_file,
             "timestamp_utc": time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime()), "device": self.device,
             "mpi_ranks": self.mpi_ranks}
        d.update(copy.deepcopy(self.sections))
        kp = d["kpoints"]
        kp["n_irreducible"] = len(kp["list"])
        return _plain(d)

    def write(self, path):
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "w", encoding="utf-8") as f:
            json.dump(self.as_dict(), f, indent=2, sort_keys=True, ensure_ascii=False)

    @classmethod
    def from_calculator(cls, calc, test_name="", input_file="", forces=None, timing=None):
        """Filled from a Calculator after run() (and compute_forces(), whose
        result ``forces`` may be passed; without it the forces section keeps
        its defaults), as tests/e2e/test_EndToEnd.cpp:312-470 fills it.
        Stress is not computed here: its section keeps ``computed`` false."""
        w = cls()
        w.test_name, w.input_file, w.device, w.mpi_ranks, w.lynx_version = test_name, input_file, "GPU", 1, "lynx_amd"
        c = calc.config
        lv = np.asarray(c["lattice"]["vectors"], float)
        orth = calc.orthogonal
        types = []
        for t, at in enumerate(c["atoms"]):
            types.append({"element": at.get("element", ""), "Zval": calc.zions[t], "mass_amu": 1.0,
                          "n_atoms": int(np.asarray(at["coordinates"]).size // 3), "pseudo_file": at["pseudo_file"]})
        # positions as the reference holds them (Crystal.cpp:283-295):
        # Cartesian in an orthogonal cell, frac_i |a_i| along the lattice
        # directions otherwise; "fractional" is Lattice::cart_to_frac of them
        pos = calc.frac @ lv if orth else calc.frac * np.linalg.norm(lv, axis=1)
        w.set_system(n_atoms=calc.natom, n_electrons=calc.Ne, atom_types=types, atoms_cartesian_bohr=pos,
                     atoms_fractional=pos @ np.linalg.inv(lv))
        w.set_cell(lattice_vectors_bohr=lv, volume_bohr3=abs(np.linalg.det(lv)), orthogonal=orth)
        g = c["grid"]
        w.set_grid(Nx=int(g["Nx"]), Ny=int(g["Ny"]), Nz=int(g["Nz"]), dx_bohr=calc.h[0], dy_bohr=calc.h[1],
                   dz_bohr=calc.h[2], fd_order=int(g.get("fd_order", 12)), boundary_conditions=["P", "P", "P"])
        w.set_electronic(xc={"GGA_PBE": "PBE"}.get(calc.xc, calc.xc), spin="collinear" if calc.nspin == 2 else "none",
                         soc=False, smearing="gaussian" if calc.smearing == 1 else "fermi-dirac",
                         electronic_temperature_K=calc.temperature_K, kBT_Ha=1.0 / calc.beta,
                         n_states=calc.nstates, cheb_degree=calc.cheb_degree)
        kp = c.get("kpoints", {})
        red, wts = calc.kpoints_reduced()
        w.set_kpoints(grid=[int(k) for k in kp.get("grid", [1, 1, 1])],
                      shift=[float(s) for s in kp.get("shift", [0, 0, 0])],
                      list=[{"index": i + 1, "reduced": red[i], "weight": wts[i]} for i in range(len(wts))])
        r = calc.energy()
        h = calc.scf.history()
        na = max(calc.natom, 1)
        s = calc.scf_kw
        w.set_scf(max_iter=s["max_iter"], tolerance=s["tol"], mixing_variable="density",
                  preconditioner="kerker" if s["kerker"] else "none", mixing_history=s["mixing_history"],
                  mixing_parameter=s["mixing_param"], converged=bool(r["converged"]), n_iterations=r["iterations"],
                  history=[{"iter": i + 1, "free_energy_per_atom_Ha": h["etotal"][i] / na,
                            "scf_error": h["error"][i], "time_s": h["seconds"][i]} for i in range(len(h["etotal"]))])
        w.set_energies(total_free_energy=r["etotal"], free_energy_per_atom=r["etotal"] / calc.natom if calc.natom else 0.0,
                       band_structure=r["eband"], exchange_correlation=r["exc"], hartree=r["ehart"],
                       self_energy=r["eself"], correction_energy=r["ec"], self_plus_correction=r["eself"] + r["ec"],
                       entropy_term=r["entropy"], fermi_level=r["fermi"])
        if forces is not None:
            f = np.asarray(forces, float).reshape(-1, 3)
            w.set_forces(max=float(np.abs(f).max()) if f.size else 0.0,
                         rms=float(math.sqrt((f * f).sum() / (3.0 * len(f)))) if len(f) else 0.0, per_atom=f)
        eig, occ, _, _ = calc.scf.state()
        ns, nk, nb = calc.nspin, len(wts), calc.nstates
        eig, occ = eig.reshape(ns, nk, nb), occ.reshape(ns, nk, nb)
        per = []
        for k in range(nk):
            e = {"kpoint_index": k + 1, "reduced": red[k], "weight": wts[k],
                 "spin_up": {"eigenvalues": eig[0, k], "occupations": occ[0, k]}}
            if ns == 2:
                e["spin_down"] = {"eigenvalues": eig[1, k], "occupations": occ[1, k]}
            per.append(e)
        w.set_eigenvalues(n_states=nb, per_kpoint=per)
        if timing:
            w.set_timing(**timing)
        return w
