"""SCF driver and the pylynx-style Calculator on the MI355X library.

``SCF`` wraps lynx_scf_* (the FP64 SCF loop on the GPU around the CheFSI
eigen stage).  ``Calculator`` mirrors pylynx's Calculator
(python/src/Calculator.hpp:22-119): ``load_config(json)`` reads the
reference's JSON input keys (src/io/InputParser.cpp:68-242), ``setup()``
builds the grid, atoms and k-points, ``run()`` runs the SCF and returns the
total energy, with the accessors ``energy()``, ``fermi_energy()``,
``n_iterations()``, ``is_converged()``, and ``compute_forces()`` after a
run.  Defaults follow
ParameterDefaults.cpp:10-80.  The "output" keys density_restart_file and
density_output_file read and write the reference's density file
(SCF.cpp:742-761, 822-831; lynx_amd.io), ``set_initial_density`` starts
from a given density, and ``write_results`` writes the reference's result
JSON.  All numerical work goes through the C ABI.
"""
from __future__ import annotations

import ctypes
import json
import math
import os

import time

import numpy as np

from . import _lib as L
from . import io

KB = 3.1668115634556e-6    # Ha / K
EH = 27.211386245988       # eV per Ha

XC = {"LDA_PZ": L.XC_LDA_PZ, "LDA_PW": L.XC_LDA_PW, "GGA_PBE": L.XC_GGA_PBE}


class ScfParams(ctypes.Structure):
    _fields_ = [("nspin", ctypes.c_int), ("nkpt", ctypes.c_int), ("complex_kpt", ctypes.c_int),
                ("nband", ctypes.c_int), ("cheb_degree", ctypes.c_int), ("rho_trigger", ctypes.c_int),
                ("nchefsi", ctypes.c_int), ("beta", ctypes.c_double), ("smearing", ctypes.c_int),
                ("xc", ctypes.c_int), ("tol", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("min_iter", ctypes.c_int), ("poisson_tol", ctypes.c_double), ("precond_tol", ctypes.c_double),
                ("mixing_history", ctypes.c_int), ("mixing_param", ctypes.c_double), ("kerker", ctypes.c_int),
                ("fp64_switch_residual", ctypes.c_double), ("low_precision", ctypes.c_int),
                ("linear_solver", ctypes.c_int), ("gram_precision", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("etotal", ctypes.c_double), ("eband", ctypes.c_double), ("exc", ctypes.c_double),
                ("ehart", ctypes.c_double), ("eself", ctypes.c_double), ("ec", ctypes.c_double),
                ("entropy", ctypes.c_double), ("fermi", ctypes.c_double), ("scf_error", ctypes.c_double),
                ("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("natom", ctypes.c_int),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SCF:
    """FP64 SCF loop on the GPU for the Hamiltonian ``H`` (set up with atoms
    from psp8 files): nband orbitals per (spin, k-point), Fermi-Dirac
    (smearing 0) or Gaussian (1) occupations at inverse temperature beta.
    ``solver`` picks how the Poisson and Kerker equations are solved:
    "spectral" (FFT, exact to rounding) or "aar" (the reference's iteration
    to poisson_tol / precond_tol).  ``gram`` ("fp32", "bf16x3" or
    "bf16x3_rot") picks the Gram products (and with "bf16x3_rot" the
    rotations) of the reduced-precision passes."""

    def __init__(self, H, nband, beta, xc="GGA_PBE", nspin=1, kpts=None, kpt_weights=None, smearing=0,
                 cheb_degree=20, rho_trigger=4, nchefsi=1, tol=1e-6, max_iter=100, min_iter=2, poisson_tol=0.0,
                 precond_tol=0.0, mixing_history=7, mixing_param=0.3, kerker=True, fp64_switch_residual=0.0,
                 low_precision="fp32", solver="spectral", gram="fp32"):
        self.H = H
        self.kpts = None if kpts is None else np.ascontiguousarray(kpts, float).reshape(-1, 3)
        nk = 1 if self.kpts is None else len(self.kpts)
        self.nspin, self.nkpt, self.nband = nspin, nk, nband
        p = ScfParams(nspin, nk, int(self.kpts is not None), nband, cheb_degree, rho_trigger, nchefsi, float(beta),
                      smearing, XC[xc] if isinstance(xc, str) else int(xc), float(tol), max_iter, min_iter,
                      float(poisson_tol), float(precond_tol), mixing_history, float(mixing_param), int(kerker),
                      float(fp64_switch_residual), {"fp32": L.FILTER_FP32, "bf16": L.FILTER_BF16}[low_precision],
                      {"spectral": L.SOLVE_SPECTRAL, "aar": L.SOLVE_AAR}[solver],
                      {"fp32": L.GRAM_FP32, "bf16x3": L.GRAM_BF16X3, "bf16x3_rot": L.GRAM_BF16X3_ROT}[gram])
        kw = None if kpt_weights is None else np.ascontiguousarray(kpt_weights, float)
        self._s = ctypes.c_void_p()
        L.check(L.lib().lynx_scf_create(ctypes.byref(self._s), H.ctx, ctypes.byref(p),
                                        None if self.kpts is None else L.dptr(self.kpts),
                                        None if kw is None else L.dptr(kw)), H.ctx)

    def _check(self, st):
        if st != L.OK:
            raise L.LynxError(st, L.lib().lynx_scf_last_error(self._s).decode())

    def run(self):
        r = ScfResult()
        self._check(L.lib().lynx_scf_run(self._s, ctypes.byref(r)))
        self.natom = r.natom
        return r.as_dict()

    def set_initial_density(self, rho, mag=None):
        """Start the next run() from the total density ``rho`` (Nd) and,
        for two spins, the magnetization ``mag`` (None: split evenly) in
        place of the atomic densities; ``rho`` None restores the atomic
        start (SCF::set_initial_density, SCF.cpp:83-107)."""
        if rho is None:
            self._check(L.lib().lynx_scf_set_initial_density(self._s, None, 0, None))
            return
        r = np.ascontiguousarray(rho, float).ravel()
        m = None if mag is None else np.ascontiguousarray(mag, float).ravel()
        if m is not None and m.size != r.size:
            raise ValueError("set_initial_density: mag and rho differ in size")
        self._keep = (r, m)
        self._check(L.lib().lynx_scf_set_initial_density(self._s, L.dptr(r), r.size,
                                                         None if m is None else L.dptr(m)))

    def history(self):
        n = ctypes.c_int()
        self._check(L.lib().lynx_scf_history(self._s, 0, None, None, None, None, ctypes.byref(n)))
        m = n.value
        e, err, t = np.zeros(m), np.zeros(m), np.zeros(m)
        prec = np.zeros(m, np.int32)
        self._check(L.lib().lynx_scf_history(self._s, m, L.dptr(e), L.dptr(err), L.dptr(t), L.iptr(prec),
                                             ctypes.byref(n)))
        return {"etotal": e, "error": err, "seconds": t, "precision": prec}

    def state(self):
        nd = self.H.grid.Nd()
        nb = self.nspin * self.nkpt * self.nband
        eig, occ = np.zeros(nb), np.zeros(nb)
        rho, veff = np.zeros(self.nspin * nd), np.zeros(self.nspin * nd)
        self._check(L.lib().lynx_scf_get(self._s, L.dptr(eig), L.dptr(occ), L.dptr(rho), L.dptr(veff)))
        return eig, occ, rho, veff

    def forces(self, with_parts=False):
        """Atomic forces (natom x 3, Ha/Bohr, Cartesian) of the last run's
        state, mean removed (Forces::compute); with_parts also returns the
        local, nonlocal and core-correction terms before the mean is
        removed."""
        n = getattr(self, "natom", 0)   # known after run(); before it the library reports LYNX_E_STATE
        f = np.zeros(3 * max(n, 1))
        parts = np.zeros(9 * max(n, 1))
        self._check(L.lib().lynx_scf_forces(self._s, L.dptr(f), L.dptr(parts)))
        f = f.reshape(n, 3)
        if with_parts:
            return f, [parts[i * 3 * n:(i + 1) * 3 * n].reshape(n, 3) for i in range(3)]
        return f

    def build_potential(self, rho):
        """XC + Poisson + Veff of a given density (nspin x Nd: total, or up and down)."""
        r = np.ascontiguousarray(rho, float)
        self._check(L.lib().lynx_scf_build_potential(self._s, L.dptr(r)))
        return self.fields()

    def fields(self):
        nd = self.H.grid.Nd()
        phi, exc, vxc = np.zeros(nd), np.zeros(nd), np.zeros(self.nspin * nd)
        self._check(L.lib().lynx_scf_get_fields(self._s, L.dptr(phi), L.dptr(exc), L.dptr(vxc)))
        _, _, _, veff = self.state()
        return {"phi": phi, "exc": exc, "vxc": vxc, "veff": veff}

    def close(self):
        if self._s:
            L.lib().lynx_scf_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- Calculator


def h_eff(h):
    """ParameterDefaults.cpp:10-18."""
    if abs(h[0] - h[1]) < 1e-12 and abs(h[1] - h[2]) < 1e-12:
        return h[0]
    return math.sqrt(3.0 / sum(1.0 / (x * x) for x in h))


def cheb_degree_default(he):
    """ParameterDefaults.cpp:20-33."""
    if he > 0.7:
        return 14
    p3, p2, p1, p0 = -700.0 / 3.0, 1240.0 / 3.0, -773.0 / 3.0, 1078.0 / 15.0
    return int(round(((p3 * he + p2) * he + p1) * he + p0))


def elec_temp_default(smearing):
    """ParameterDefaults.cpp:35-39: 0.2 eV (Gaussian) or 0.1 eV (Fermi-Dirac)."""
    ev = 0.2 if smearing == "gaussian" else 0.1
    return 1.0 / (KB * (EH / ev))


class Calculator:
    """pylynx's Calculator workflow on the MI355X library: the reference's
    JSON input, grid, atoms (psp8), k-points and the SCF."""

    def __init__(self, device_id=0):
        self.device_id = device_id
        self.config = None
        self._dir = "."
        self._setup = False
        self._result = None
        self._rho0 = None
        self._forces = None
        self._times = {"scf_total": 0.0, "force": 0.0, "stress": 0.0, "walltime_total": 0.0}
        self._t0 = time.perf_counter()

    def load_config(self, json_file):
        with open(json_file) as f:
            self.config = json.load(f)
        self._dir = os.path.dirname(os.path.abspath(json_file))
        return self

    def set_config(self, config, base_dir="."):
        self.config = config
        self._dir = base_dir
        return self

    def setup(self):
        from . import core
        c = self.config
        latvec = np.asarray(c["lattice"]["vectors"], float)
        orth = c["lattice"].get("cell_type", "orthogonal") == "orthogonal"
        g = c["grid"]
        n = (int(g["Nx"]), int(g["Ny"]), int(g["Nz"]))
        order = int(g.get("fd_order", 12))
        lat = core.Lattice(latvec, core.CellType.Orthogonal if orth else core.CellType.NonOrthogonal)
        grid = core.FDGrid(*n, lat)
        st = core.FDStencil(order, grid, lat)
        dom = core.Domain(grid)
        files, types, frac = [], [], []
        for t, at in enumerate(c["atoms"]):
            pf = at["pseudo_file"]
            files.append(pf if os.path.isabs(pf) else os.path.join(self._dir, pf))
            coords = np.asarray(at["coordinates"], float).reshape(-1, 3)
            if not at.get("fractional", True):
                coords = coords @ np.linalg.inv(latvec)
            for x in coords:
                types.append(t)
                frac.append(x)
        self.frac = np.array(frac)
        vnl = core.NonlocalProjector().setup_atoms(files, types, self.frac, True)
        self.H = core.Hamiltonian(self.device_id).setup(st, dom, grid, core.HaloExchange(dom, st.FDn()), vnl)
        self.orthogonal = orth
        self.shape = n
        self.latvec = latvec
        self.h = (grid.dx(), grid.dy(), grid.dz())
        self.dV = grid.dV()
        zions = []
        for f in files:
            z = ctypes.c_double(); lm = ctypes.c_int(); ll = ctypes.c_int(); npj = ctypes.c_int()
            rc = ctypes.c_double()
            L.check(L.lib().lynx_psp8_info(f.encode(), ctypes.byref(z), ctypes.byref(lm), ctypes.byref(ll),
                                           ctypes.byref(npj), ctypes.byref(rc)))
            zions.append(z.value)
        self.zions = zions
        self.Ne = int(sum(int(zions[t]) for t in types))
        self.natom = len(types)
        e = c.get("electronic", {})
        self.xc = e.get("xc", "GGA_PBE")
        self.nspin = 2 if e.get("spin", "none") in ("collinear", "polarized") else 1
        smear = e.get("smearing", "gaussian")
        temp = float(e.get("temperature", -1.0))
        if temp <= 0:
            temp = elec_temp_default(smear)
        self.temperature_K = temp
        self.beta = 1.0 / (KB * temp)
        self.smearing = 1 if smear == "gaussian" else 0
        self.nstates = int(e.get("Nstates", 0)) or (self.Ne // 2 + 10)
        s = c.get("scf", {})
        h = (grid.dx(), grid.dy(), grid.dz())
        deg = int(s.get("cheb_degree", -1))
        self.cheb_degree = deg if deg > 0 else cheb_degree_default(h_eff(h))
        self.scf_kw = dict(tol=float(s.get("tolerance", 1e-6)), max_iter=int(s.get("max_iter", 100)),
                           min_iter=int(s.get("min_iter", 2)), mixing_history=int(s.get("mixing_history", 7)),
                           mixing_param=float(s.get("mixing_parameter", 0.3)),
                           kerker=s.get("preconditioner", "kerker") == "kerker",
                           rho_trigger=int(s.get("rho_trigger", 4)), nchefsi=int(s.get("nchefsi", 1)))
        if s.get("mixing", "density") != "density":
            raise NotImplementedError("potential mixing is not built; density mixing only")
        kp = c.get("kpoints")
        self.kpts = self.kw = None
        if kp and any(int(k) > 1 for k in kp.get("grid", [1, 1, 1])):
            K = [int(k) for k in kp["grid"]]
            sh = np.ascontiguousarray(kp.get("shift", [0, 0, 0]), float)
            lv = np.ascontiguousarray(latvec, float)
            cell = L.CELL_ORTHOGONAL if orth else L.CELL_NONORTHOGONAL
            nk = ctypes.c_int()
            L.check(L.lib().lynx_kpoints(K[0], K[1], K[2], L.dptr(sh), L.dptr(lv), cell, None, None, None, 0,
                                         ctypes.byref(nk)))
            kc, w = np.zeros(3 * nk.value), np.zeros(nk.value)
            L.check(L.lib().lynx_kpoints(K[0], K[1], K[2], L.dptr(sh), L.dptr(lv), cell, L.dptr(kc), None, L.dptr(w),
                                         nk.value, ctypes.byref(nk)))
            kr = np.zeros(3 * nk.value)
            L.check(L.lib().lynx_kpoints(K[0], K[1], K[2], L.dptr(sh), L.dptr(lv), cell, None, L.dptr(kr), None,
                                         nk.value, ctypes.byref(nk)))
            self.kpts, self.kw, self.kred = kc.reshape(-1, 3), w, kr.reshape(-1, 3)
        self._setup = True
        return self

    def run(self, **overrides):
        if not self._setup:
            self.setup()
        kw = dict(self.scf_kw)
        kw.update(overrides)
        self.scf = SCF(self.H, self.nstates, self.beta, xc=self.xc, nspin=self.nspin, kpts=self.kpts,
                       kpt_weights=self.kw, smearing=self.smearing, cheb_degree=self.cheb_degree, **kw)
        out = self.config.get("output", {})
        rho0 = self._rho0
        restart = out.get("density_restart_file")
        if restart:
            # SCF.cpp:742-761: a file that does not fit falls back to the
            # atomic densities with a warning
            path = restart if os.path.isabs(restart) else os.path.join(self._dir, restart)
            got = io.read_density(path, self.shape, self.nspin)
            if got is None:
                print(f"WARNING: Failed to read density from {restart}, using atomic density", flush=True)
            else:
                rho0 = io.total_and_mag(got)
                print(f"Density restart: loaded from {restart} (Ne={got.sum() * self.dV:.6f})", flush=True)
        if rho0 is not None:
            self.scf.set_initial_density(*rho0)
        t = time.perf_counter()
        self._result = self.scf.run()
        self._times["scf_total"] = time.perf_counter() - t
        self._forces = None
        dest = out.get("density_output_file")
        if dest:
            path = dest if os.path.isabs(dest) else os.path.join(self._dir, dest)
            if io.write_density(path, self.spin_densities(), self.shape, self.latvec, self.dV):
                print(f"Density written to {dest}", flush=True)
            else:
                print(f"WARNING: Failed to write density to {dest}", flush=True)
        return self._result["etotal"]

    def compute_forces(self):
        """Atomic forces after run(), (Natom, 3) in Ha/Bohr (pylynx's
        Calculator.compute_forces, python/src/bind_calculator.cpp:104-112)."""
        if self._result is None:
            raise RuntimeError("compute_forces needs a completed run()")
        t = time.perf_counter()
        self._forces = self.scf.forces()
        self._times["force"] = time.perf_counter() - t
        return self._forces

    def set_initial_density(self, rho, mag=None):
        """Start the next run() from the total density ``rho`` (Nd) and the
        magnetization ``mag`` (two spins; None splits evenly) instead of the
        atomic densities (python/src/bind_calculator.cpp:123-139)."""
        self._rho0 = None if rho is None else (np.array(rho, float).ravel(),
                                               None if mag is None else np.array(mag, float).ravel())

    def spin_densities(self):
        """The density the last run ended on, nspin x Nd (total for one
        spin, up and down for two): the input density of its last iteration,
        the one the reference writes to its density file."""
        _, _, rho, _ = self.scf.state()
        return rho.reshape(self.nspin, -1)

    def density(self):
        """Total density of the last run, Nd (pylynx's Calculator.density,
        python/src/bind_calculator.cpp:177-179)."""
        r = self.spin_densities()
        return r[0] if self.nspin == 1 else r[0] + r[1]

    def kpoints_reduced(self):
        """Reduced coordinates and weights of the k-points in use."""
        if self.kpts is None:
            return np.zeros((1, 3)), np.ones(1)
        return self.kred, self.kw

    def write_results(self, path, test_name="", input_file=""):
        """The reference's result JSON of the last run (lynx_amd.io.ResultWriter)."""
        if self._result is None:
            raise RuntimeError("write_results needs a completed run()")
        self._times["walltime_total"] = time.perf_counter() - self._t0
        io.ResultWriter.from_calculator(self, test_name=test_name, input_file=input_file, forces=self._forces,
                                        timing=self._times).write(path)

    # accessors of pylynx's Calculator
    def is_setup(self): return self._setup
    def is_converged(self): return bool(self._result and self._result["converged"])
    def energy(self): return self._result
    def fermi_energy(self): return self._result["fermi"]
    def n_iterations(self): return self._result["iterations"]
    def Natom(self): return self.natom
    def Nelectron(self): return self.Ne
    def Nspin(self): return self.nspin
