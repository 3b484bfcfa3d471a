"""Host-side mirror of the pylynx ``_core`` classes on the CheFSI path.

Class, method and argument names follow LYNX's pybind11 module
(python/src/bind_core.cpp:20-144, bind_operators.cpp:51-78,
bind_solvers.cpp:54-114) so that code written against ``lynx._core`` runs
unchanged on the MI355X library.  All numerical work goes through the C ABI
(``liblynx_chefsi.so``); nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes
from enum import IntEnum

import numpy as np

from . import _lib as L


class CellType(IntEnum):
    Orthogonal = 0
    NonOrthogonal = 1


class BCType(IntEnum):
    Periodic = 0
    Dirichlet = 1


class Lattice:
    """Lattice(latvec, cell_type) -- rows of latvec are the lattice vectors."""

    def __init__(self, latvec, cell_type=CellType.Orthogonal):
        self._latvec = np.ascontiguousarray(latvec, dtype=np.float64).reshape(3, 3)
        self._cell_type = CellType(cell_type)

    def latvec(self):
        return self._latvec.copy()

    def cell_type(self):
        return self._cell_type

    def is_orthogonal(self):
        return self._cell_type == CellType.Orthogonal

    def lengths(self):
        return np.sqrt((self._latvec ** 2).sum(axis=1))

    def jacobian(self):
        return abs(np.linalg.det(self._latvec))

    def __repr__(self):
        return f"Lattice(lengths={tuple(np.round(self.lengths(), 6))}, {self._cell_type.name})"


class FDGrid:
    """FDGrid(Nx, Ny, Nz, lattice, bcx, bcy, bcz); periodic boundaries only."""

    def __init__(self, Nx, Ny, Nz, lattice, bcx=BCType.Periodic, bcy=BCType.Periodic, bcz=BCType.Periodic):
        if any(BCType(b) != BCType.Periodic for b in (bcx, bcy, bcz)):
            raise ValueError("the CheFSI path is periodic in all directions")
        self._n = (int(Nx), int(Ny), int(Nz))
        self._lattice = lattice

    def Nx(self): return self._n[0]
    def Ny(self): return self._n[1]
    def Nz(self): return self._n[2]
    def Nd(self): return self._n[0] * self._n[1] * self._n[2]

    def dx(self): return self._lattice.lengths()[0] / self._n[0]
    def dy(self): return self._lattice.lengths()[1] / self._n[1]
    def dz(self): return self._lattice.lengths()[2] / self._n[2]

    def dV(self): return self._lattice.jacobian() / self.Nd()

    def lattice(self): return self._lattice

    def __repr__(self):
        return f"FDGrid({self._n[0]}x{self._n[1]}x{self._n[2]})"


class FDStencil:
    """FDStencil(order, grid, lattice)."""

    def __init__(self, order, grid, lattice):
        self._order = int(order)
        self.grid = grid
        self.lattice = lattice

    def order(self): return self._order
    def FDn(self): return self._order // 2

    def weights(self):
        n = self.FDn() + 1
        w1 = np.zeros(n); w2 = np.zeros(n)
        L.check(L.lib().lynx_fd_weights(self._order, L.dptr(w1), L.dptr(w2)))
        return w1, w2


class Domain:
    """Domain(grid, verts): the whole grid (no domain decomposition)."""

    def __init__(self, grid, verts=None):
        self.grid = grid

    def Nx_d(self): return self.grid.Nx()
    def Ny_d(self): return self.grid.Ny()
    def Nz_d(self): return self.grid.Nz()
    def Nd_d(self): return self.grid.Nd()


def full_domain(grid):
    return Domain(grid)


class HaloExchange:
    """HaloExchange(domain, FDn).  The wrap is fused into the stencil kernel,
    so this object only carries the extended sizes."""

    def __init__(self, domain, FDn):
        self.domain = domain
        self._fdn = int(FDn)

    def FDn(self): return self._fdn
    def nx_ex(self): return self.domain.Nx_d() + 2 * self._fdn
    def ny_ex(self): return self.domain.Ny_d() + 2 * self._fdn
    def nz_ex(self): return self.domain.Nz_d() + 2 * self._fdn
    def nd_ex(self): return self.nx_ex() * self.ny_ex() * self.nz_ex()


class NonlocalProjector:
    """Kleinman-Bylander projectors of a crystal.

    ``setup_atoms(psp_files, atom_types, coords, fractional=True)`` reads the
    psp8 files and builds the projectors in the library (C++ host code);
    ``setup_raw(...)`` takes projectors built elsewhere."""

    def __init__(self):
        self._atoms = None
        self._raw = None

    def setup_atoms(self, psp_files, atom_types, coords, fractional=True):
        self._atoms = (list(psp_files), np.ascontiguousarray(atom_types, np.int32),
                       np.ascontiguousarray(coords, np.float64).reshape(-1, 3), bool(fractional))
        self._raw = None
        return self

    def setup_raw(self, img_atom, img_ndc, img_nproj, gpos, chi, img_shift, n_atom, ip_displ, gamma):
        self._raw = dict(img_atom=np.ascontiguousarray(img_atom, np.int32),
                         img_ndc=np.ascontiguousarray(img_ndc, np.int32),
                         img_nproj=np.ascontiguousarray(img_nproj, np.int32),
                         gpos=np.ascontiguousarray(gpos, np.int32),
                         chi=np.ascontiguousarray(chi, np.float64),
                         img_shift=np.ascontiguousarray(img_shift, np.float64),
                         n_atom=int(n_atom),
                         ip_displ=np.ascontiguousarray(ip_displ, np.int32),
                         gamma=np.ascontiguousarray(gamma, np.float64))
        self._atoms = None
        return self

    def is_setup(self):
        return self._atoms is not None or self._raw is not None

    def _install(self, ctx):
        lib = L.lib()
        if self._atoms is not None:
            files, types, coords, frac = self._atoms
            arr = (ctypes.c_char_p * len(files))(*[f.encode() for f in files])
            L.check(lib.lynx_set_atoms(ctx, len(files), arr, len(types), L.iptr(types), L.dptr(coords),
                                       int(frac)), ctx)
        elif self._raw is not None:
            r = self._raw
            L.check(lib.lynx_set_nonlocal(ctx, len(r["img_atom"]), L.iptr(r["img_atom"]), L.iptr(r["img_ndc"]),
                                          L.iptr(r["img_nproj"]), L.iptr(r["gpos"]), L.dptr(r["chi"]),
                                          L.dptr(r["img_shift"]), r["n_atom"], L.iptr(r["ip_displ"]),
                                          L.dptr(r["gamma"])), ctx)


class _Context:
    """Owns one lynx_ctx (HIP stream, device buffers, operator data)."""

    def __init__(self, device_id=0):
        p = ctypes.c_void_p()
        L.check(L.lib().lynx_ctx_create(ctypes.byref(p), int(device_id)))
        self.ptr = p

    def __del__(self):
        try:
            if self.ptr:
                L.lib().lynx_ctx_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass


class Hamiltonian:
    """H = -1/2 lap + Veff + V_nl on the GPU.

    ``setup(stencil, domain, grid, halo, vnl=None)`` and
    ``apply(halo, psi, Veff, c=0.0)`` follow bind_operators.cpp:51-71; psi is
    an (Nd,) or (Nd, ncol) array whose memory is taken as ncol consecutive
    columns of Nd values, exactly as the reference wrapper passes it
    (OperatorWrappers.hpp:98-119).  ``apply_kpt`` takes complex columns."""

    def __init__(self, device_id=0):
        self._c = _Context(device_id)
        self._veff_key = None
        self._veff_gen = -1
        self.stencil = None
        self._kpt = (0.0, 0.0, 0.0)

    @property
    def ctx(self):
        return self._c.ptr

    def setup(self, stencil, domain, grid, halo, vnl=None):
        lat = grid.lattice()
        lv = np.ascontiguousarray(lat.latvec(), np.float64)
        L.check(L.lib().lynx_set_grid(self.ctx, grid.Nx(), grid.Ny(), grid.Nz(), L.dptr(lv),
                                      int(lat.cell_type()), stencil.order()), self.ctx)
        self.stencil, self.domain, self.grid, self.halo = stencil, domain, grid, halo
        self.vnl = vnl
        self._veff_key = None
        if vnl is not None and vnl.is_setup():
            vnl._install(self.ctx)
        return self

    def set_veff(self, Veff):
        v = np.ascontiguousarray(Veff, np.float64).ravel()
        if v.size != self.grid.Nd():
            raise ValueError("Veff must have Nd entries")
        # skip the upload only when the values are bit-identical to the last
        # ones this object uploaded (the caller may change an array in place)
        # and no other caller has set a potential on the context since (the
        # eigen stage sets each spin's Veff itself)
        last = self._veff_key
        if (last is not None and self._veff_gen == self.veff_generation() and last.size == v.size
                and np.array_equal(last.view(np.int64), v.view(np.int64))):
            return
        # a refused potential leaves the context without one: forget the
        # previous upload first so that it is sent again next time
        self._veff_key = None
        L.check(L.lib().lynx_set_veff(self.ctx, L.dptr(v), L.MEM_HOST), self.ctx)
        self._veff_key = v.copy()
        self._veff_gen = self.veff_generation()

    def veff_generation(self):
        g = ctypes.c_long()
        L.check(L.lib().lynx_get_veff_generation(self.ctx, ctypes.byref(g)), self.ctx)
        return g.value

    def invalidate_veff(self):
        """Forget the cached potential so that the next call uploads it."""
        self._veff_key = None

    def set_kpoint(self, kpt_cart):
        k = np.ascontiguousarray(kpt_cart, np.float64)
        L.check(L.lib().lynx_set_kpoint(self.ctx, L.dptr(k)), self.ctx)
        self._kpt = tuple(k)

    def grid_info(self):
        h = np.zeros(3); dv = ctypes.c_double(); me = ctypes.c_double()
        L.check(L.lib().lynx_get_grid_info(self.ctx, L.dptr(h), ctypes.byref(dv), ctypes.byref(me)), self.ctx)
        return h, dv.value, me.value

    def _apply(self, psi, Veff, c, dtype):
        self.set_veff(Veff)
        nd = self.grid.Nd()
        x = np.ascontiguousarray(psi, dtype=dtype)
        if x.size % nd:
            raise ValueError("psi size must be a multiple of Nd")
        ncol = x.size // nd
        y = np.empty_like(x)
        L.check(L.lib().lynx_h_apply(self.ctx, L.dtype_code(dtype), L.vptr(x), nd, L.vptr(y), nd, ncol,
                                     float(c), L.MEM_HOST), self.ctx)
        return y

    def apply(self, halo, psi, Veff, c=0.0):
        return self._apply(psi, Veff, c, np.float64)

    def apply_kpt(self, halo, psi, Veff, kpt_cart, c=0.0):
        self.set_kpoint(kpt_cart)
        return self._apply(psi, Veff, c, np.complex128)


class Wavefunction:
    """Orbitals, eigenvalues and occupations per (spin, k-point), with the
    pylynx Wavefunction interface (bind_electronic.cpp:15-44).  ``psi`` blocks
    are (Nd_d, Nband) Fortran-ordered arrays; ``randomize`` fills one from
    glibc's srand(seed)/rand() stream through lynx_randomize
    (Wavefunction.cpp:53-76)."""

    def __init__(self):
        self._nd = self._nband = self._nspin = self._nkpts = 0
        self._complex = False
        self._psi, self._eig, self._occ = [], [], []

    def allocate(self, Nd_d, Nband, Nspin, Nkpts, is_complex=False, Nspinor=1):
        if Nspinor != 1:
            raise ValueError("spinor wavefunctions are outside the CheFSI path")
        self._nd, self._nband, self._nspin, self._nkpts = int(Nd_d), int(Nband), int(Nspin), int(Nkpts)
        self._complex = bool(is_complex)
        dt = np.complex128 if self._complex else np.float64
        n = self._nspin * self._nkpts
        self._psi = [np.zeros((self._nd, self._nband), dt, order="F") for _ in range(n)]
        self._eig = [np.zeros(self._nband) for _ in range(n)]
        self._occ = [np.zeros(self._nband) for _ in range(n)]
        return self

    def Nd_d(self): return self._nd
    def Nband(self): return self._nband
    def Nband_global(self): return self._nband
    def Nspin(self): return self._nspin
    def Nkpts(self): return self._nkpts
    def is_complex(self): return self._complex

    def _i(self, spin, kpt):
        if not (0 <= spin < self._nspin and 0 <= kpt < self._nkpts):
            raise IndexError("spin or k-point out of range")
        return spin * self._nkpts + kpt

    def randomize(self, spin, kpt, seed=42):
        X = self._psi[self._i(spin, kpt)]
        L.check(L.lib().lynx_randomize(None, L.dtype_code(X.dtype), L.vptr(X), self._nd, self._nd, self._nband,
                                       int(seed), L.MEM_HOST))

    def psi(self, spin=0, kpt=0): return self._psi[self._i(spin, kpt)]
    def eigenvalues(self, spin=0, kpt=0): return self._eig[self._i(spin, kpt)]
    def occupations(self, spin=0, kpt=0): return self._occ[self._i(spin, kpt)]


class LynxContext:
    """The parts of LynxContext the eigensolver reads: grid objects, the halo
    and the number of states (LynxContext.hpp)."""

    def __init__(self, lattice, grid, stencil, domain, halo, Nstates=0):
        self._lattice, self._grid, self._stencil = lattice, grid, stencil
        self._domain, self._halo, self._Nstates = domain, halo, int(Nstates)

    def lattice(self): return self._lattice
    def grid(self): return self._grid
    def stencil(self): return self._stencil
    def domain(self): return self._domain
    def halo(self): return self._halo
    def Nstates(self): return self._Nstates


class EigenSolver:
    """CheFSI on the GPU, with the pylynx EigenSolver interface
    (bind_solvers.cpp:54-114)."""

    def __init__(self):
        self.H = None
        self._lambda_cutoff = 0.0
        self._nband = 0

    def setup(self, ctx, H):
        self.H = H
        self._nband = ctx.Nstates() if ctx is not None else 0
        return self

    def _solve(self, psi, Veff, lambda_cutoff, eigval_min, eigval_max, cheb_degree, dtype):
        if self.H is None:
            raise RuntimeError("EigenSolver.setup(ctx, H) has not been called")
        H = self.H
        H.set_veff(Veff)
        psi_out = np.array(psi, dtype=dtype, order="F", copy=True)
        if psi_out.ndim != 2:
            raise ValueError("psi must be 2D (Nd_d, Nband)")
        nd, nband = psi_out.shape
        if nd != H.grid.Nd():
            raise ValueError("psi rows must equal Nd")
        eig = np.zeros(nband)
        lc = ctypes.c_double()
        L.check(L.lib().lynx_chefsi_solve(H.ctx, L.dtype_code(dtype), L.vptr(psi_out), nd, nband, L.dptr(eig),
                                          float(lambda_cutoff), float(eigval_min), float(eigval_max),
                                          int(cheb_degree), ctypes.byref(lc), L.MEM_HOST), H.ctx)
        self._lambda_cutoff = lc.value
        return psi_out, eig

    def solve(self, psi, Veff, lambda_cutoff, eigval_min, eigval_max, cheb_degree=20):
        """One CheFSI pass (EigenSolver.cpp:738-832).  Returns (psi, eigvals)."""
        return self._solve(psi, Veff, lambda_cutoff, eigval_min, eigval_max, cheb_degree, np.float64)

    def solve_kpt(self, psi, Veff, lambda_cutoff, eigval_min, eigval_max, kpt_cart, cell_lengths=None,
                  cheb_degree=20):
        """Complex Bloch solve (EigenSolver.cpp:834-842)."""
        self.H.set_kpoint(kpt_cart)
        return self._solve(psi, Veff, lambda_cutoff, eigval_min, eigval_max, cheb_degree, np.complex128)

    def lanczos_bounds(self, Veff, tol=1e-2, max_iter=1000):
        """Spectral bounds (Lanczos.hpp:104-198), seed 1 as the reference."""
        return self._lanczos(Veff, tol, max_iter, L.F64)

    def lanczos_bounds_kpt(self, Veff, kpt_cart, cell_lengths=None, tol=1e-2, max_iter=1000):
        self.H.set_kpoint(kpt_cart)
        return self._lanczos(Veff, tol, max_iter, L.C128)

    def _lanczos(self, Veff, tol, max_iter, code):
        self.H.set_veff(Veff)
        mn = ctypes.c_double(); mx = ctypes.c_double(); it = ctypes.c_int()
        L.check(L.lib().lynx_lanczos_bounds(self.H.ctx, code, ctypes.byref(mn), ctypes.byref(mx), float(tol),
                                            int(max_iter), 1, ctypes.byref(it)), self.H.ctx)
        self.lanczos_iters = it.value
        return mn.value, mx.value

    def lambda_cutoff(self): return self._lambda_cutoff
    def set_lambda_cutoff(self, v): self._lambda_cutoff = float(v)
    def Nband_global(self): return self._nband
    def is_band_parallel(self): return False


class BatchFilter:
    """Chebyshev filter of many independent units in one pass: unit i is
    ``(H, kpt)`` with a set-up Hamiltonian (its grid, atoms and Veff) and a
    k-point in the lattice-aligned basis (``None`` for Gamma).  The k-points
    of one structure share one Hamiltonian; structures with different grids
    mix freely.  Each Chebyshev step is one launch per kernel for all units,
    replayed from a HIP graph (lynx_batch_filter)."""

    def __init__(self, units, graph=True):
        self.units = list(units)
        n = len(self.units)
        ctxs = (ctypes.c_void_p * n)(*[H.ctx.value for H, _ in self.units])
        k = np.zeros(3 * n)
        for i, (_, kpt) in enumerate(self.units):
            if kpt is not None:
                k[3 * i:3 * i + 3] = kpt
        self._b = ctypes.c_void_p()
        st = L.lib().lynx_batch_create(ctypes.byref(self._b), n, ctxs, L.dptr(k))
        if st != L.OK:
            raise L.LynxError(st, "lynx_batch_create failed: %s" % L.lib().lynx_status_string(st).decode())
        self._check(L.lib().lynx_batch_set_graph(self._b, 1 if graph else 0))

    def _check(self, st):
        if st != L.OK:
            raise L.LynxError(st, L.lib().lynx_batch_last_error(self._b).decode())

    def filter(self, blocks, degree, lambda_cutoff, eig_min, eig_max, dtype=None):
        """blocks: list of (Nd, ncol) arrays (column-major data); returns the
        filtered blocks in the same dtype (float64, complex128, float32 or
        complex64)."""
        n = len(self.units)
        if dtype is None:
            dtype = blocks[0].dtype
        xs = [np.asfortranarray(b, dtype=dtype) for b in blocks]
        ys = [np.zeros_like(x) for x in xs]
        xp = (ctypes.c_void_p * n)(*[x.ctypes.data for x in xs])
        yp = (ctypes.c_void_p * n)(*[y.ctypes.data for y in ys])
        ld = (ctypes.c_long * n)(*[x.shape[0] for x in xs])
        nc = np.array([x.shape[1] for x in xs], np.int32)
        lc = np.broadcast_to(np.asarray(lambda_cutoff, float), (n,)).copy()
        lo = np.broadcast_to(np.asarray(eig_min, float), (n,)).copy()
        hi = np.broadcast_to(np.asarray(eig_max, float), (n,)).copy()
        self._check(L.lib().lynx_batch_filter(self._b, L.dtype_code(dtype), xp, ld, yp, ld, L.iptr(nc), int(degree),
                                              L.dptr(lc), L.dptr(lo), L.dptr(hi), L.MEM_HOST))
        return ys

    def filter_device(self, dtype, xptrs, yptrs, lds, ncols, degree, lambda_cutoff, eig_min, eig_max):
        """Device-resident variant: lists of device addresses and sizes."""
        n = len(self.units)
        xp = (ctypes.c_void_p * n)(*xptrs)
        yp = (ctypes.c_void_p * n)(*yptrs)
        ld = (ctypes.c_long * n)(*lds)
        nc = np.asarray(ncols, np.int32)
        lc = np.broadcast_to(np.asarray(lambda_cutoff, float), (n,)).copy()
        lo = np.broadcast_to(np.asarray(eig_min, float), (n,)).copy()
        hi = np.broadcast_to(np.asarray(eig_max, float), (n,)).copy()
        self._check(L.lib().lynx_batch_filter(self._b, L.dtype_code(dtype), xp, ld, yp, ld, L.iptr(nc), int(degree),
                                              L.dptr(lc), L.dptr(lo), L.dptr(hi), L.MEM_DEVICE))

    def set_filter_precision(self, precision):
        """"fp64" (default) or "fp32": the filter of solve_device on
        single-precision copies (lynx_batch_set_filter_precision)."""
        self._check(L.lib().lynx_batch_set_filter_precision(
            self._b, {"fp64": L.FILTER_FP64, "fp32": L.FILTER_FP32}[precision]))

    def solve_device(self, dtype, ptrs, lds, ncols, degree, lambda_cutoff, eig_min, eig_max):
        """One CheFSI pass for every unit (lynx_batch_solve) on device blocks
        (lists of addresses, leading dimensions and column counts), in
        float64 or complex128; returns (eigenvalues per unit, lambda_cutoff
        per unit)."""
        n = len(self.units)
        pp = (ctypes.c_void_p * n)(*ptrs)
        ld = (ctypes.c_long * n)(*lds)
        nc = np.asarray(ncols, np.int32)
        lc = np.broadcast_to(np.asarray(lambda_cutoff, float), (n,)).copy()
        lo = np.broadcast_to(np.asarray(eig_min, float), (n,)).copy()
        hi = np.broadcast_to(np.asarray(eig_max, float), (n,)).copy()
        eig = np.zeros(int(nc.sum()))
        lco = np.zeros(n)
        self._check(L.lib().lynx_batch_solve(self._b, L.dtype_code(dtype), pp, ld, L.iptr(nc), L.dptr(eig),
                                             L.dptr(lc), L.dptr(lo), L.dptr(hi), int(degree), L.dptr(lco)))
        off = np.concatenate([[0], np.cumsum(nc)])
        return [eig[off[i]:off[i + 1]] for i in range(n)], lco

    def last_timing(self):
        ms = ctypes.c_double(); nl = ctypes.c_int()
        self._check(L.lib().lynx_batch_last_timing(self._b, ctypes.byref(ms), ctypes.byref(nl)))
        return ms.value, nl.value

    def close(self):
        if self._b:
            L.lib().lynx_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ScfEigParams(ctypes.Structure):
    _fields_ = [("nspin", ctypes.c_int), ("nkpt", ctypes.c_int), ("complex_kpt", ctypes.c_int),
                ("nband", ctypes.c_int), ("cheb_degree", ctypes.c_int), ("rho_trigger", ctypes.c_int),
                ("nchefsi", ctypes.c_int), ("nelectron", ctypes.c_double), ("beta", ctypes.c_double),
                ("smearing", ctypes.c_int), ("fp64_switch_residual", ctypes.c_double),
                ("low_precision", ctypes.c_int), ("gram_precision", ctypes.c_int)]


class ScfEigenStage:
    """The eigen stage of an SCF iteration (SCF::solve_eigenproblem with its
    spectral-bound bookkeeping) on one Hamiltonian: ``step(scf_iter,
    residual, Veff, psis)`` runs the CheFSI passes over every (spin, k-point)
    block and returns eigenvalues, occupations and the Fermi level.  With
    ``fp64_switch_residual`` > 0 the pass runs in reduced precision
    (``low_precision``: "fp32", or "bf16" for bfloat16 filter storage) until
    the SCF residual drops below it, with the Gram products of those passes
    as FP32 GEMMs or split-BF16x3 MFMA products (``gram``: "fp32" or
    "bf16x3"; "bf16x3_rot" also rotates that way) (lynx_scf_eig_*)."""

    def __init__(self, H, nband, nelectron, beta, kpts=None, kpt_weights=None, nspin=1, cheb_degree=20,
                 rho_trigger=4, nchefsi=1, smearing=0, fp64_switch_residual=0.0, low_precision="fp32", gram="fp32"):
        self.H = H
        self.kpts = None if kpts is None else np.ascontiguousarray(kpts, float).reshape(-1, 3)
        nk = 1 if self.kpts is None else len(self.kpts)
        self.nspin, self.nkpt, self.nband = nspin, nk, nband
        self.complex_ = self.kpts is not None
        p = _ScfEigParams(nspin, nk, int(self.complex_), nband, cheb_degree, rho_trigger, nchefsi,
                          float(nelectron), float(beta), smearing, float(fp64_switch_residual),
                          {"fp32": L.FILTER_FP32, "bf16": L.FILTER_BF16}[low_precision],
                          {"fp32": L.GRAM_FP32, "bf16x3": L.GRAM_BF16X3, "bf16x3_rot": L.GRAM_BF16X3_ROT}[gram])
        kw = None if kpt_weights is None else np.ascontiguousarray(kpt_weights, float)
        self._d = ctypes.c_void_p()
        L.check(L.lib().lynx_scf_eig_create(ctypes.byref(self._d), H.ctx, ctypes.byref(p),
                                            None if self.kpts is None else L.dptr(self.kpts),
                                            None if kw is None else L.dptr(kw)), H.ctx)

    def step(self, scf_iter, residual, Veff, psis):
        """psis: list of nspin*nkpt (Nd, nband) Fortran-ordered arrays, updated in place."""
        dt = np.complex128 if self.complex_ else np.float64
        for p in psis:
            if p.dtype != dt or not p.flags.f_contiguous:
                raise ValueError("psi blocks must be Fortran-ordered %s arrays" % np.dtype(dt).name)
        n = len(psis)
        ptrs = (ctypes.c_void_p * n)(*[p.ctypes.data for p in psis])
        v = np.ascontiguousarray(Veff, np.float64)
        eig = np.zeros(n * self.nband)
        occ = np.zeros(n * self.nband)
        ef = ctypes.c_double()
        prec = ctypes.c_int()
        st = L.lib().lynx_scf_eig_step(self._d, int(scf_iter), float(residual), L.dptr(v), ptrs, psis[0].shape[0],
                                       L.MEM_HOST, L.dptr(eig), L.dptr(occ), ctypes.byref(ef), ctypes.byref(prec))
        if st != L.OK:
            raise L.LynxError(st, L.lib().lynx_scf_eig_last_error(self._d).decode())
        return eig, occ, ef.value, prec.value

    def step_device(self, scf_iter, residual, Veff, ptrs, ld):
        """Device-resident blocks: ptrs are device addresses (lynx_amd.hipmem)."""
        n = len(ptrs)
        pp = (ctypes.c_void_p * n)(*ptrs)
        v = np.ascontiguousarray(Veff, np.float64)
        eig = np.zeros(n * self.nband)
        occ = np.zeros(n * self.nband)
        ef = ctypes.c_double()
        prec = ctypes.c_int()
        st = L.lib().lynx_scf_eig_step(self._d, int(scf_iter), float(residual), L.dptr(v), pp, int(ld),
                                       L.MEM_DEVICE, L.dptr(eig), L.dptr(occ), ctypes.byref(ef), ctypes.byref(prec))
        if st != L.OK:
            raise L.LynxError(st, L.lib().lynx_scf_eig_last_error(self._d).decode())
        return eig, occ, ef.value, prec.value

    def state(self):
        lc = ctypes.c_double()
        mn = np.zeros(self.nspin); mx = np.zeros(self.nspin)
        L.check(L.lib().lynx_scf_eig_state(self._d, ctypes.byref(lc), L.dptr(mn), L.dptr(mx)))
        return lc.value, mn, mx

    def __del__(self):
        try:
            if self._d:
                L.lib().lynx_scf_eig_destroy(self._d)
        except Exception:
            pass
