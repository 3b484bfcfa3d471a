"""Synthetic test systems shared by the tests, smoke() and bench.py.

Cells follow the workloads the CheFSI path is measured on: diamond Si (C.pot
stands in for the absent Si psp8: same zion 4, lmax 1, lloc 4, two s and two
p projectors), displaced supercells, Gaussian-well effective potentials.
Random numbers come from numpy's PCG64 (seeded) unless a glibc rand() stream
is asked for.
"""
from __future__ import annotations

import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PSP_DIR = os.path.join(ROOT, "tests", "golden", "psp")

DIAMOND = np.array([[0.00, 0.00, 0.00], [0.25, 0.25, 0.25], [0.50, 0.50, 0.00], [0.75, 0.75, 0.25],
                    [0.50, 0.00, 0.50], [0.75, 0.25, 0.75], [0.00, 0.50, 0.50], [0.25, 0.75, 0.75]])
A_SI = 10.26


def psp(name="C.pot"):
    return os.path.join(PSP_DIR, name)


def supercell(n, base=DIAMOND):
    """n x n x n supercell of the 8-atom cubic cell, fractional coordinates."""
    out = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                out.append((base + np.array([i, j, k])) / n)
    return np.concatenate(out)


def displace(frac, a, sigma, seed):
    """Gaussian displacement sigma (Bohr) per Cartesian component."""
    rng = np.random.default_rng(seed)
    return frac + rng.normal(0.0, sigma, frac.shape) / a


FCC_PRIMITIVE = 0.5 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
SI2 = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])


def gaussian_veff_cell(n, latvec, frac, depth=1.2, width=1.6, offset=0.3):
    """Gaussian wells at every atom of a general cell (rows of latvec are
    the lattice vectors), nearest periodic image in fractional coordinates."""
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    g = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    V = np.full(g.shape[0], offset)
    for f in np.asarray(frac):
        df = g - f
        df -= np.round(df)
        d = df @ latvec
        V -= depth * np.exp(-(d * d).sum(axis=1) / (2 * width ** 2))
    return V


def gaussian_veff(n, a, frac, depth=1.2, width=1.6, offset=0.3):
    """Effective potential with a Gaussian well at every atom (periodic
    minimum image), a stand-in for the atomic-superposition Veff."""
    nx, ny, nz = n
    h = np.array([a / nx, a / ny, a / nz])
    x = np.arange(nx) * h[0]
    y = np.arange(ny) * h[1]
    z = np.arange(nz) * h[2]
    V = np.full((nz, ny, nx), offset)
    for f in np.asarray(frac) * a:
        dx = (x - f[0] + a / 2) % a - a / 2
        dy = (y - f[1] + a / 2) % a - a / 2
        dz = (z - f[2] + a / 2) % a - a / 2
        r2 = dz[:, None, None] ** 2 + dy[None, :, None] ** 2 + dx[None, None, :] ** 2
        V -= depth * np.exp(-r2 / (2 * width ** 2))
    return V.ravel()


class System:
    """One cubic cell: grid, atoms, Veff, and both the oracle operator and
    (optionally) the GPU Hamiltonian built from the same inputs."""

    def __init__(self, n, a, frac, psp_name="C.pot", order=12, with_nonlocal=True, kpt=(0.0, 0.0, 0.0),
                 latvec=None, psp_names=None, types=None, extra_wells=None):
        if isinstance(n, int):
            n = (n, n, n)
        self.n = tuple(n)
        self.a = a
        self.frac = np.asarray(frac, float)
        self.order = order
        self.orth = latvec is None
        if latvec is None:
            L = np.broadcast_to(np.asarray(a, float), (3,))
            self.latvec = np.diag(L).astype(float)
        else:
            self.latvec = np.asarray(latvec, float)
        self.psp_files = [psp(p) for p in (psp_names or [psp_name])]
        self.types = np.zeros(len(self.frac), np.int32) if types is None else np.asarray(types, np.int32)
        self.with_nonlocal = with_nonlocal and len(self.frac) > 0
        wells = self.frac if extra_wells is None else np.concatenate([self.frac, np.asarray(extra_wells, float)])
        if self.orth and np.ndim(a) == 0:
            self.veff = gaussian_veff(self.n, a, wells)
        else:
            self.veff = gaussian_veff_cell(self.n, self.latvec, wells)
        self.nd = int(np.prod(self.n))
        self.kpt = tuple(kpt)

    # -- oracle ---------------------------------------------------------------
    def oracle_op(self):
        import oracle
        lat = oracle.Lattice(self.latvec, self.orth)
        g = oracle.Grid(*self.n, lat)
        if self.with_nonlocal:
            ps = [oracle.load_psp8(f) for f in self.psp_files]
            return oracle.setup_system(g, ps, list(self.types), self.frac, self.order, self.kpt)
        return oracle.Operator(g, self.order, kpt=self.kpt)

    # -- GPU library -------------------------------------------------------------
    def hamiltonian(self, device_id=0):
        from lynx_amd import core
        lat = core.Lattice(self.latvec, core.CellType.Orthogonal if self.orth else core.CellType.NonOrthogonal)
        grid = core.FDGrid(*self.n, lat)
        st = core.FDStencil(self.order, grid, lat)
        dom = core.Domain(grid)
        halo = core.HaloExchange(dom, st.FDn())
        vnl = None
        if self.with_nonlocal:
            vnl = core.NonlocalProjector().setup_atoms(self.psp_files, self.types, self.frac, True)
        H = core.Hamiltonian(device_id).setup(st, dom, grid, halo, vnl)
        if any(abs(k) > 0 for k in self.kpt):
            H.set_kpoint(self.kpt)
        self.core = (lat, grid, st, dom, halo)
        return H

    def context(self, nstates):
        from lynx_amd import core
        lat, grid, st, dom, halo = self.core
        return core.LynxContext(lat, grid, st, dom, halo, nstates)

    def random_block(self, ncol, seed, complex_=False):
        rng = np.random.default_rng(seed)
        X = rng.uniform(-0.5, 0.5, (self.nd, ncol))
        if complex_:
            X = X + 1j * rng.uniform(-0.5, 0.5, (self.nd, ncol))
        return np.asfortranarray(X)


def si8(n=30, **kw):
    return System(n, A_SI, DIAMOND, **kw)


def si16(seed, sigma=0.05):
    """Si16: the Si8 cell doubled along z (10.26 x 10.26 x 20.52 Bohr), 30 x 30 x 60."""
    base = np.concatenate([DIAMOND * [1, 1, 0.5], (DIAMOND + [0, 0, 1]) * [1, 1, 0.5]])
    rng = np.random.default_rng(seed)
    L = np.array([A_SI, A_SI, 2 * A_SI])
    frac = base + rng.normal(0.0, sigma, base.shape) / L
    return System((30, 30, 60), L, frac)


def batio3(seed, n=28, a=7.57, sigma=0.05):
    """Cubic BaTiO3 (a = 7.57 Bohr), Ti displaced 0.1 Bohr along z.  No Ba or
    Ti psp8 ships with the reference; La.pot and Zr.pot (its test-suite
    neighbours of Ba and Ti) stand in, with O.pot for oxygen."""
    frac = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5 + 0.1 / a],
                     [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    rng = np.random.default_rng(seed)
    frac = frac + rng.normal(0.0, sigma, frac.shape) / a
    return System(n, a, frac, psp_names=["La.pot", "Zr.pot", "O.pot"], types=[0, 1, 2, 2, 2])


def h2o(seed, n=50, box=20.0, sigma=0.05):
    """H2O in a 20 Bohr box.  The reference ships no H psp8: the two H atoms
    enter only through Veff wells; O carries O.pot projectors."""
    o = np.array([0.5, 0.5, 0.5])
    hs = o + np.array([[1.431, 1.108, 0.0], [-1.431, 1.108, 0.0]]) / box
    rng = np.random.default_rng(seed)
    o = o + rng.normal(0.0, sigma, 3) / box
    hs = hs + rng.normal(0.0, sigma, hs.shape) / box
    return System(n, box, o[None, :], psp_names=["O.pot"], extra_wells=hs)


def si2_fcc(n=24, **kw):
    """Si2 in the FCC primitive cell (non-orthogonal), a = 10.26 Bohr."""
    return System(n, A_SI, SI2, latvec=A_SI * FCC_PRIMITIVE, **kw)


# -- the measured workloads ------------------------------------------------------

A_SI64, A_SI216 = 2 * A_SI, 30.78


def si64_filter_bench():
    """Si64 filter-only operator workload: the 2x2x2 diamond supercell
    (a = 20.52 Bohr), atoms displaced by sigma = 0.05 Bohr (seed 1), 60^3 grid."""
    return System(60, A_SI64, displace(supercell(2), A_SI64, 0.05, seed=1))


def si216_mixed():
    """Si216: the 3x3x3 supercell (a = 30.78 Bohr), sigma = 0.03 Bohr (seed 3), 90^3 grid."""
    return System(90, A_SI216, displace(supercell(3), A_SI216, 0.03, seed=3))


def hetero_frames():
    """The 64-frame heterogeneous batch: (class, system, orbitals) for 32 Si8
    (30^3, 24), 16 Si16 (30x30x60, 40), 8 BaTiO3 (28^3, 30), 8 H2O (50^3, 8),
    sigma = 0.05 Bohr (seeds 4 + frame)."""
    frames = [("Si8", System(30, A_SI, displace(DIAMOND, A_SI, 0.05, seed=4 + i)), 24) for i in range(32)]
    frames += [("Si16", si16(36 + i), 40) for i in range(16)]
    frames += [("BaTiO3", batio3(52 + i), 30) for i in range(8)]
    frames += [("H2O", h2o(60 + i), 8) for i in range(8)]
    return frames


def fcc_kpoints(s, k=(6, 6, 6)):
    """Monkhorst-Pack k-points of a cell, reduced by time reversal
    (lynx_kpoints), in the lattice-aligned basis the library takes."""
    import ctypes
    from lynx_amd import _lib as L
    lv = np.ascontiguousarray(s.latvec, float)
    cell = L.CELL_ORTHOGONAL if s.orth else L.CELL_NONORTHOGONAL
    n = ctypes.c_int()
    sh = np.zeros(3)
    L.check(L.lib().lynx_kpoints(k[0], k[1], k[2], L.dptr(sh), L.dptr(lv), cell, None, None, None, 0,
                                 ctypes.byref(n)))
    kc = np.zeros(3 * n.value)
    w = np.zeros(n.value)
    L.check(L.lib().lynx_kpoints(k[0], k[1], k[2], L.dptr(sh), L.dptr(lv), cell, L.dptr(kc), None, L.dptr(w),
                                 n.value, ctypes.byref(n)))
    return kc.reshape(-1, 3), w
