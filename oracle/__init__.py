"""CPU oracle for the CheFSI hot path -- TEST INFRASTRUCTURE ONLY.

This package restates LYNX's algorithm on the CPU so that the HIP path can be
checked against it.  It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by ``bench.py``'s ``cpu_baseline`` leg, and
there only as the checker.  The product library never loads it.

Layout:
  * numpy restatements of the setup steps (lattice, FD weights, psp8 reader,
    Hermite splines, real spherical harmonics, nonlocal influence spheres,
    Chi projectors, Monkhorst-Pack k-points);
  * ctypes bindings to ``oracle.c`` (``liboracle.so``) for the grid-sized
    loops: halo, stencil, nonlocal apply, Chebyshev filter, Lanczos;
  * numpy/scipy LAPACK for the dense subspace steps (Cholesky, triangular
    solve, symmetric eigensolve), the same textbook routines the reference
    calls through BLAS/LAPACK.

Each function names the reference file:line it follows.  Pinning status of
each function (what independent check fixes it) is listed in ``PINNING``.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# What fixes each oracle function independently of its own formula.
PINNING = {
    "fd_weights": "exact order-6 fractions and sum rule from tests/unit/test_FDStencil.cpp:7-105",
    "stencil_coeffs": "D2 = w2/h^2 scaling (test_FDStencil.cpp:59-75); plane-wave symbol closed form",
    "max_eigval_half_lap": "closed form of the plane-wave symbol at the Nyquist vector",
    "halo": "wrap and Bloch-phase values of tests/unit/test_HaloExchange.cpp:12-219",
    "h_apply (local)": "-1/2 lap sin x = 1/2 sin x (test_Hamiltonian.cpp:39-72); plane-wave symbol; H const = V const",
    "h_apply (nonlocal)": "dense brute force V_nl = dV sum Chi G Chi^T; symmetry / hermiticity",
    "load_psp8": "file values at the nodes; r*p(r)/r with r=0 copy (Pseudopotential.cpp:143-157)",
    "spline": "exact for cubics; nodal interpolation",
    "ylm_real": "orthonormality on the sphere and the addition theorem",
    "chi": "addition theorem makes V_nl independent of the m convention; sphere membership by brute force",
    "cheb_filter": "closed form p_m(lambda) = T_m((l-c)/e)/T_m((lmin-c)/e) on eigenvectors",
    "lanczos": "the reference's own lanczos_bounds<T> (Lanczos.hpp, LAPACK dsterf_) run on this oracle's H.v: "
               "same step count and bounds; bounds bracket the dense spectrum; glibc rand() start vector",
    "rand_uniform": "the reference's Wavefunction::randomize[_kpt], compiled from source, bitwise",
    "fermi_level": "the reference's Occupation::compute, compiled from source, on metallic and gapped spectra "
                   "(in a gap the root is flat and the Brent steps fix the point); closed forms (symmetric "
                   "levels, one half-filled level, sum w f = Ne)",
    "ScfEigStage": "composed of the pinned solve, Lanczos and Fermi level; its bookkeeping by properties a "
                   "plausible mistake fails: a constant Veff shift of one spin shifts that spin's Lanczos bounds "
                   "exactly and makes it set lambda_c; eig_min kept at the Lanczos value in iteration 0 and "
                   "taken from eig(s, k0)[0] afterwards; eig_max fixed; solve count rho_trigger / nchefsi "
                   "passes x spins x k-points; sum w f = Ne",
    "solve": "dense eigh of the explicit H on a small grid (subspace converges to lowest N)",
    "kpoints": "MP 6x6x6 -> 112 after time reversal (KPoints.cpp:55-70); weights sum",
}


def lib():
    """Load liboracle.so, building it with the Makefile if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


# ---------------------------------------------------------------------------
# Lattice (src/core/Lattice.cpp:13-54) and grid spacing (src/core/FDGrid.cpp:6-35)

@dataclass
class Lattice:
    latvec: np.ndarray           # rows are lattice vectors (Bohr)
    orthogonal: bool

    def __post_init__(self):
        self.latvec = np.asarray(self.latvec, dtype=np.float64).reshape(3, 3)
        self.lengths = np.sqrt((self.latvec ** 2).sum(axis=1))
        self.uvec = self.latvec / self.lengths[:, None]
        self.jacobian = abs(np.linalg.det(self.latvec))
        self.uvec_inv = np.linalg.inv(self.uvec)
        # lapc_T = (U^-1)^T U^-1 (Lattice.cpp:41-42)
        self.lapc_T = self.uvec_inv.T @ self.uvec_inv
        # sphere extent factors ||column_i(U^-1)|| (Lattice.cpp:44-53)
        self.nc_extent = np.sqrt((self.uvec_inv ** 2).sum(axis=0))

    def noncart_to_cart(self, nc):
        """Lattice.cpp:105-115: cart_j = sum_i nc_i U(i, j)."""
        return np.asarray(nc) @ self.uvec


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    lattice: Lattice

    def __post_init__(self):
        L = self.lattice.lengths
        self.h = np.array([L[0] / self.nx, L[1] / self.ny, L[2] / self.nz])
        self.dV = self.lattice.jacobian / (self.nx * self.ny * self.nz)
        self.nd = self.nx * self.ny * self.nz


# ---------------------------------------------------------------------------
# FD stencil (src/operators/FDStencil.cpp)

def _fract(n, k):
    """FDStencil.cpp:7-14: n! / ((n-k)! (n+k)!) as a running product."""
    num = 1.0
    for i in range(n - k + 1, n + 1):
        num *= i
    den = 1.0
    for i in range(n + 1, n + k + 1):
        den *= i
    return num / den


def fd_weights(order):
    """FDStencil.cpp:23-44: central weights of the first and second derivative."""
    fdn = order // 2
    w1 = np.zeros(fdn + 1)
    w2 = np.zeros(fdn + 1)
    for p in range(1, fdn + 1):
        sgn = 1.0 if p % 2 == 1 else -1.0
        w1[p] = sgn * _fract(fdn, p) / p
        w2[0] -= 2.0 / (p * p)
        w2[p] = sgn * 2.0 * _fract(fdn, p) / (p * p)
    return w1, w2


@dataclass
class Stencil:
    fdn: int
    D1: np.ndarray        # (3, fdn+1)
    D2: np.ndarray        # (3, fdn+1): x, y, z
    D2m: np.ndarray       # (3, fdn+1): xy, xz, yz mixed coefficients
    max_eig: float


def stencil_coeffs(order, grid: Grid):
    """FDStencil.cpp:46-130: scaled coefficients and the -1/2 Laplacian bound."""
    w1, w2 = fd_weights(order)
    fdn = order // 2
    h = grid.h
    lat = grid.lattice
    D1 = np.stack([w1 / h[a] for a in range(3)])
    if lat.orthogonal:
        D2 = np.stack([w2 / (h[a] * h[a]) for a in range(3)])
        D2m = np.zeros((3, fdn + 1))
    else:
        T = lat.lapc_T
        D2 = np.stack([T[a, a] * w2 / (h[a] * h[a]) for a in range(3)])
        D2m = np.stack([2.0 * T[0, 1] * w1 / h[0],
                        2.0 * T[0, 2] * w1 / h[0],
                        2.0 * T[1, 2] * w1 / h[1]])
    if lat.orthogonal:
        n = (grid.nx, grid.ny, grid.nz)
        scal = [(n[a] - n[a] % 2) / n[a] for a in range(3)]
        m = D2[0, 0] + D2[1, 0] + D2[2, 0]
        for p in range(1, fdn + 1):
            m += 2.0 * sum(D2[a, p] * math.cos(math.pi * p * scal[a]) for a in range(3))
        max_eig = -0.5 * m
    else:
        m = abs(D2[0, 0]) + abs(D2[1, 0]) + abs(D2[2, 0])
        for p in range(1, fdn + 1):
            m += 2.0 * (abs(D2[0, p]) + abs(D2[1, p]) + abs(D2[2, p]))
            m += 2.0 * (abs(D2m[0, p]) + abs(D2m[1, p]) + abs(D2m[2, p]))
        max_eig = 0.5 * m
    return Stencil(fdn, D1, D2, D2m, max_eig)


# ---------------------------------------------------------------------------
# psp8 reader (src/atoms/Pseudopotential.cpp:28-333) and Hermite splines
# (src/core/NumericalMethods.cpp:128-225)

@dataclass
class Psp8:
    zion: float
    pspxc: int
    lmax: int
    lloc: int
    mmax: int
    ppl: list
    r: np.ndarray
    udv: dict = field(default_factory=dict)       # (l, p) -> p_l(r) on r
    udv_d: dict = field(default_factory=dict)     # spline first derivatives
    gamma: dict = field(default_factory=dict)     # (l, p) -> ekb
    rc: list = field(default_factory=list)
    r_vloc: np.ndarray | None = None
    vloc0: float = 0.0                             # Vloc(r = 0) as listed in the file
    fchrg: float = 0.0                             # model core charge switch
    rho_c: np.ndarray | None = None                # core density / (4 pi)
    rho_iso: np.ndarray | None = None              # isolated-atom valence density / (4 pi)

    @property
    def nproj_per_atom(self):
        """Pseudopotential.cpp:19-26."""
        return sum(self.ppl[l] * (2 * l + 1) for l in range(self.lmax + 1) if l != self.lloc)

    @property
    def rc_max(self):
        return max(self.rc) if self.rc else 0.0


def _floats(line):
    """Leading numeric fields of a line (stream extraction stops at the first
    non-number, as the reference's istringstream reads do); Fortran 'D'
    exponents are accepted."""
    out = []
    for t in line.replace("D", "E").replace("d", "E").split():
        try:
            out.append(float(t))
        except ValueError:
            break
    return out


def load_psp8(path):
    """Pseudopotential.cpp:28-333 (nonlocal part plus the local channel)."""
    with open(path) as f:
        lines = f.read().splitlines()
    pos = 1
    zion = _floats(lines[pos])[1]; pos += 1
    t = lines[pos].split(); pos += 1
    pspcod, pspxc, lmax, lloc, mmax = (int(t[0]), int(t[1]), int(t[2]), int(t[3]), int(t[4]))
    if pspcod != 8:
        raise ValueError("only psp8 is supported")
    fchrg = _floats(lines[pos])[1]; pos += 1    # rchrg fchrg qchrg
    nums = lines[pos].split(); pos += 1
    ppl = [int(nums[l]) for l in range(lmax + 1)]
    ext = int(lines[pos].split()[0]); pos += 1
    nprojso = [0] * (lmax + 1)
    if ext >= 2:
        so = _floats(lines[pos]); pos += 1       # nprojso line (l = 1..lmax)
        for l in range(1, lmax + 1):
            if l - 1 < len(so):
                nprojso[l] = int(so[l - 1])
    r = np.zeros(mmax)
    psp = Psp8(zion, pspxc, lmax, lloc, mmax, ppl, r)
    psp.fchrg = fchrg
    have_r = False
    for l in range(lmax + 1):
        hdr = _floats(lines[pos]); pos += 1
        nprj = ppl[l]
        if l != lloc and nprj > 0:
            for p in range(nprj):
                psp.gamma[(l, p)] = hdr[1 + p]
            rows = np.array([_floats(lines[pos + i]) for i in range(mmax)]); pos += mmax
            ri = rows[:, 1]
            if not have_r:
                r[:] = ri
            for p in range(nprj):
                v = rows[:, 2 + p]
                u = np.where(ri > 1e-10, v / np.where(ri > 1e-10, ri, 1.0), 0.0)
                if mmax > 1:
                    u[0] = u[1]
                psp.udv[(l, p)] = u
            have_r = True
        elif l == lloc:
            rows = np.array([_floats(lines[pos + i]) for i in range(mmax)]); pos += mmax
            if not have_r:
                r[:] = rows[:, 1]
            psp.r_vloc = rows[:, 1] * rows[:, 2]
            psp.vloc0 = rows[0, 2]
            have_r = True
    if lloc > lmax:
        pos += 1
        rows = np.array([_floats(lines[pos + i]) for i in range(mmax)]); pos += mmax
        if not have_r:
            r[:] = rows[:, 1]
        psp.r_vloc = rows[:, 1] * rows[:, 2]
        psp.vloc0 = rows[0, 2]
    if psp.vloc0 == 0.0 and mmax > 1:
        psp.vloc0 = psp.r_vloc[1] / r[1] if r[0] < 1e-14 else psp.r_vloc[0] / r[0]
    # spin-orbit projectors (skipped), model core charge, isolated-atom
    # valence density: "i r value ..." rows, value / (4 pi)
    if ext >= 2 and lmax >= 1:
        for l in range(1, lmax + 1):
            if nprojso[l] > 0:
                pos += 1 + mmax

    def density_block():
        nonlocal pos
        vals = []
        while len(vals) < mmax and pos < len(lines):
            f = _floats(lines[pos])
            if len(f) < 3:
                break
            vals.append(f[2] / (4.0 * math.pi))
            pos += 1
        return np.array(vals + [0.0] * (mmax - len(vals))) if vals else None

    if fchrg > 1e-10:
        psp.rho_c = density_block()
    psp.rho_iso = density_block()
    psp.r_vloc_d = spline_deriv(r, psp.r_vloc)
    psp.rho_c_d = None if psp.rho_c is None else spline_deriv(r, psp.rho_c)
    psp.rho_iso_d = None if psp.rho_iso is None else spline_deriv(r, psp.rho_iso)
    for l in range(lmax + 1):
        if ppl[l] > 0 and l != lloc:
            u = psp.udv[(l, 0)]
            nz = np.nonzero(np.abs(u) > 1e-12)[0]
            psp.rc.append(float(r[nz[-1]]) if len(nz) else float(r[-1]))
        else:
            psp.rc.append(0.0)
    for key, u in psp.udv.items():
        psp.udv_d[key] = spline_deriv(r, u)
    return psp


def spline_deriv(x, y):
    """NumericalMethods.cpp:128-192: first derivatives of the Hermite cubic
    with the reference's end conditions, solved by a forward sweep."""
    x = np.asarray(x, float); y = np.asarray(y, float)
    n = len(x)
    yd = np.zeros(n)
    if n < 2:
        return yd
    if n == 2:
        yd[:] = (y[1] - y[0]) / (x[1] - x[0])
        return yd
    A = np.zeros(n); B = np.zeros(n); C = np.zeros(n)
    h0 = x[1] - x[0]; h1 = x[2] - x[1]
    r0 = (y[1] - y[0]) / h0; r1 = (y[2] - y[1]) / h1
    B[0] = h1 * (h0 + h1); C[0] = (h0 + h1) ** 2
    yd[0] = r0 * (3 * h0 * h1 + 2 * h1 * h1) + r1 * h0 * h0
    h0v = x[1:-1] - x[:-2]; h1v = x[2:] - x[1:-1]
    r0v = (y[1:-1] - y[:-2]) / h0v; r1v = (y[2:] - y[1:-1]) / h1v
    A[1:-1] = h1v; B[1:-1] = 2 * (h0v + h1v); C[1:-1] = h0v
    yd[1:-1] = 3 * (r0v * h1v + r1v * h0v)
    h0 = x[-2] - x[-3]; h1 = x[-1] - x[-2]
    r0 = (y[-2] - y[-3]) / h0; r1 = (y[-1] - y[-2]) / h1
    A[-1] = (h0 + h1) ** 2; B[-1] = h0 * (h0 + h1)
    yd[-1] = r0 * h1 * h1 + r1 * (3 * h0 * h1 + 2 * h0 * h0)
    F = np.zeros(n)
    b = B[0]
    yd[0] /= b
    for j in range(1, n):
        F[j] = C[j - 1] / b
        b = B[j] - A[j] * F[j]
        if abs(b) < 1e-30:
            b = 1e-30
        yd[j] = (yd[j] - A[j] * yd[j - 1]) / b
    for j in range(n - 2, -1, -1):
        yd[j] -= F[j + 1] * yd[j + 1]
    return yd


def spline_eval(x, y, yd, xi):
    """NumericalMethods.cpp:194-225 / 227-: Hermite cubic on [x_j, x_j+1],
    clamped outside the table; uniform-grid interval lookup."""
    x = np.asarray(x); xi = np.asarray(xi, float)
    n = len(x)
    out = np.empty_like(xi)
    lo = xi <= x[0]
    hi = xi >= x[-1]
    dx0 = x[1] - x[0]
    uniform = np.all(np.abs(np.diff(x) - dx0) <= 1e-10 * max(1.0, dx0))
    mid = ~(lo | hi)
    r = xi[mid]
    if uniform:
        j = ((r - x[0]) / dx0).astype(np.int64)
        j = np.minimum(j, n - 2)
    else:
        j = np.searchsorted(x, r, side="right") - 1
        j = np.clip(j, 0, n - 2)
    p1 = x[j]; p3 = x[j + 1]
    idx = 1.0 / (p3 - p1)
    dy = (y[j + 1] - y[j]) * idx
    a0 = y[j]; a1 = yd[j]
    a2 = idx * (3 * dy - 2 * yd[j] - yd[j + 1])
    a3 = idx * idx * (-2 * dy + yd[j] + yd[j + 1])
    t = r - p1
    out[mid] = ((a3 * t + a2) * t + a1) * t + a0
    out[lo] = y[0]
    out[hi] = y[-1]
    return out


# ---------------------------------------------------------------------------
# Real spherical harmonics (src/core/math_utils.hpp:174-): standard real
# Y_lm with m<0 ~ sin, m>0 ~ cos; l = 0..3 written from the closed forms.

def ylm_real(l, m, x, y, z, r):
    x = np.asarray(x, float); y = np.asarray(y, float); z = np.asarray(z, float)
    r = np.asarray(r, float)
    if l == 0:
        return np.full_like(r, 0.5 * math.sqrt(1.0 / math.pi))
    small = r < 1e-14
    rr = np.where(small, 1.0, r)
    xn, yn, zn = x / rr, y / rr, z / rr
    if l == 1:
        c = math.sqrt(3.0 / (4.0 * math.pi))
        v = {-1: c * yn, 0: c * zn, 1: c * xn}[m]
    elif l == 2:
        c1 = 0.5 * math.sqrt(15.0 / math.pi)
        c0 = 0.25 * math.sqrt(5.0 / math.pi)
        c2 = 0.25 * math.sqrt(15.0 / math.pi)
        v = {-2: c1 * xn * yn, -1: c1 * yn * zn, 0: c0 * (3 * zn * zn - 1.0),
             1: c1 * xn * zn, 2: c2 * (xn * xn - yn * yn)}[m]
    elif l == 3:
        c3 = 0.25 * math.sqrt(35.0 / (2.0 * math.pi))
        c2 = 0.5 * math.sqrt(105.0 / math.pi)
        c1 = 0.25 * math.sqrt(21.0 / (2.0 * math.pi))
        c0 = 0.25 * math.sqrt(7.0 / math.pi)
        c2b = 0.25 * math.sqrt(105.0 / math.pi)
        x2, y2, z2 = xn * xn, yn * yn, zn * zn
        v = {-3: c3 * yn * (3 * x2 - y2), -2: c2 * xn * yn * zn,
             -1: c1 * yn * (5 * z2 - 1.0), 0: c0 * zn * (5 * z2 - 3.0),
             1: c1 * xn * (5 * z2 - 1.0), 2: c2b * zn * (x2 - y2),
             3: c3 * xn * (x2 - 3 * y2)}[m]
    else:
        raise NotImplementedError("l > 3")
    return np.where(small, 0.0, v)


# ---------------------------------------------------------------------------
# Atoms, influence spheres and Chi (src/atoms/Crystal.cpp:136-255,
# src/operators/NonlocalProjector.cpp:13-119)

@dataclass
class Image:
    atom: int                  # physical atom index (type-major order)
    type: int
    gpos: np.ndarray           # flat grid indices in the rc sphere
    shift: np.ndarray          # periodic image shift (lattice-aligned coords)
    chi: np.ndarray            # (ndc, nproj) projectors


def positions_noncart(lattice: Lattice, frac):
    """Crystal.cpp:284-295: fractional -> lattice-aligned (non-Cartesian)
    coordinates, x_nc = frac * |a_i|."""
    frac = np.asarray(frac, float).reshape(-1, 3)
    if lattice.orthogonal:
        return frac @ lattice.latvec
    return frac * lattice.lengths[None, :]


def nloc_influence(grid: Grid, psps, types, pos_nc):
    """Crystal.cpp:136-255: images whose rc sphere touches the grid, in the
    reference order (type, atom, iz, iy, ix), with the grid points inside."""
    lat = grid.lattice
    L = lat.lengths
    h = grid.h
    out = []
    for it, psp in enumerate(psps):
        rc = psp.rc_max
        if rc < 1e-14:
            continue
        ext = np.ones(3) if lat.orthogonal else lat.nc_extent
        for ia in range(len(types)):
            if types[ia] != it:
                continue
            pos = pos_nc[ia]
            nimg = [int(math.ceil(rc * ext[a] / L[a])) for a in range(3)]
            for iz in range(-nimg[2], nimg[2] + 1):
                for iy in range(-nimg[1], nimg[1] + 1):
                    for ix in range(-nimg[0], nimg[0] + 1):
                        img = pos + np.array([ix, iy, iz]) * L
                        if lat.orthogonal:
                            s = [int(math.ceil((img[a] - rc) / h[a])) for a in range(3)]
                            e = [int(math.floor((img[a] + rc) / h[a])) for a in range(3)]
                        else:
                            rcx = [rc * ext[a] / h[a] for a in range(3)]
                            s = [int(math.ceil(img[a] / h[a] - rcx[a])) for a in range(3)]
                            e = [int(math.floor(img[a] / h[a] + rcx[a])) for a in range(3)]
                        n = (grid.nx, grid.ny, grid.nz)
                        s = [max(s[a], 0) for a in range(3)]
                        e = [min(e[a], n[a] - 1) for a in range(3)]
                        if s[0] > e[0] or s[1] > e[1] or s[2] > e[2]:
                            continue
                        kk, jj, ii = np.meshgrid(np.arange(s[2], e[2] + 1),
                                                 np.arange(s[1], e[1] + 1),
                                                 np.arange(s[0], e[0] + 1), indexing="ij")
                        ii = ii.ravel(); jj = jj.ravel(); kk = kk.ravel()
                        d = np.stack([ii * h[0] - img[0], jj * h[1] - img[1], kk * h[2] - img[2]], axis=1)
                        if not lat.orthogonal:
                            d = lat.noncart_to_cart(d)
                        r2 = (d * d).sum(axis=1)
                        keep = r2 <= rc * rc
                        if not np.any(keep):
                            continue
                        g = (ii + jj * grid.nx + kk * grid.nx * grid.ny)[keep]
                        out.append(Image(ia, it, g.astype(np.int64),
                                         np.array([ix * L[0], iy * L[1], iz * L[2]]), None))
    return out


def build_chi(grid: Grid, psps, images, pos_nc):
    """NonlocalProjector.cpp:40-102: Chi[g, (l, p, m)] = Y_lm(r) * UdV_lp(|r|),
    columns ordered l -> p -> m = -l..l, l != lloc."""
    lat = grid.lattice
    h = grid.h
    for im in images:
        psp = psps[im.type]
        g = im.gpos
        i = g % grid.nx; j = (g // grid.nx) % grid.ny; k = g // (grid.nx * grid.ny)
        img = pos_nc[im.atom] + im.shift
        d = np.stack([i * h[0] - img[0], j * h[1] - img[1], k * h[2] - img[2]], axis=1)
        if not lat.orthogonal:
            d = lat.noncart_to_cart(d)
        r = np.sqrt((d * d).sum(axis=1))
        cols = []
        for l in range(psp.lmax + 1):
            if l == psp.lloc:
                continue
            for p in range(psp.ppl[l]):
                u = spline_eval(psp.r, psp.udv[(l, p)], psp.udv_d[(l, p)], r)
                for m in range(-l, l + 1):
                    cols.append(ylm_real(l, m, d[:, 0], d[:, 1], d[:, 2], r) * u)
        im.chi = np.stack(cols, axis=1) if cols else np.zeros((len(g), 0))
    return images


def atom_gamma(psps, types):
    """Gamma per global projector, atoms in type-major order
    (NonlocalProjector.cpp:105-112, 177-197)."""
    ip = [0]
    gam = []
    for ia, it in enumerate(types):
        psp = psps[it]
        for l in range(psp.lmax + 1):
            if l == psp.lloc:
                continue
            for p in range(psp.ppl[l]):
                gam += [psp.gamma[(l, p)]] * (2 * l + 1)
        ip.append(len(gam))
    return np.array(ip, dtype=np.int32), np.array(gam)


# ---------------------------------------------------------------------------
# Monkhorst-Pack k-points with time reversal (src/core/KPoints.cpp:8-86)

def kpoints(K, shift, lattice: Lattice):
    L = lattice.lengths
    tol = 1e-12
    s = [-int(math.floor((K[a] - 1) / 2.0)) for a in range(3)]
    sums = 2.0 * math.pi / L
    cart, red, w = [], [], []
    for n1 in range(s[0], s[0] + K[0]):
        for n2 in range(s[1], s[1] + K[1]):
            for n3 in range(s[2], s[2] + K[2]):
                kr = []
                for a, nn in enumerate((n1, n2, n3)):
                    v = nn * 1.0 / K[a]
                    v = math.fmod(v + shift[a] / K[a] + 0.5 - tol, 1.0) - 0.5 + tol
                    kr.append(v)
                kc = [kr[a] * 2.0 * math.pi / L[a] for a in range(3)]
                hit = -1
                for q, c in enumerate(cart):
                    ok = all(abs(kc[a] + c[a]) < tol or abs(kc[a] + c[a] - sums[a]) < tol for a in range(3))
                    if ok:
                        hit = q
                        break
                if hit >= 0:
                    w[hit] = 2.0
                else:
                    cart.append(kc); red.append(kr); w.append(1.0)
    w = np.array(w) / (K[0] * K[1] * K[2])
    return np.array(cart), np.array(red), w


# ---------------------------------------------------------------------------
# ctypes operator

class _OrcOp(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("fdn", ctypes.c_int), ("orth", ctypes.c_int),
        ("cx", ctypes.c_void_p), ("cy", ctypes.c_void_p), ("cz", ctypes.c_void_p),
        ("cxy", ctypes.c_void_p), ("cxz", ctypes.c_void_p), ("cyz", ctypes.c_void_p),
        ("d1y", ctypes.c_void_p), ("d1z", ctypes.c_void_p),
        ("n_img", ctypes.c_int),
        ("img_atom", ctypes.c_void_p), ("img_ndc", ctypes.c_void_p),
        ("img_gpos_off", ctypes.c_void_p), ("gpos", ctypes.c_void_p),
        ("img_chi_off", ctypes.c_void_p), ("img_nproj", ctypes.c_void_p),
        ("chi", ctypes.c_void_p), ("img_shift", ctypes.c_void_p),
        ("n_atom", ctypes.c_int), ("ip_displ", ctypes.c_void_p), ("gamma", ctypes.c_void_p),
        ("dV", ctypes.c_double),
        ("kpt", ctypes.c_double * 3), ("cell", ctypes.c_double * 3),
    ]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Operator:
    """H = -1/2 lap + Veff + V_nl on one grid; optional k-point.

    The grid-sized loops run in oracle.c; dense subspace steps use numpy.
    """

    def __init__(self, grid: Grid, order=12, images=None, n_atom=0, ip_displ=None,
                 gamma=None, kpt=(0.0, 0.0, 0.0)):
        self.grid = grid
        self.st = stencil_coeffs(order, grid)
        self.keep = []
        o = _OrcOp()
        o.nx, o.ny, o.nz = grid.nx, grid.ny, grid.nz
        o.fdn = self.st.fdn
        o.orth = 1 if grid.lattice.orthogonal else 0
        arrs = {}
        for name, a in (("cx", self.st.D2[0]), ("cy", self.st.D2[1]), ("cz", self.st.D2[2]),
                        ("cxy", self.st.D2m[0]), ("cxz", self.st.D2m[1]), ("cyz", self.st.D2m[2]),
                        ("d1y", self.st.D1[1]), ("d1z", self.st.D1[2])):
            arrs[name] = np.ascontiguousarray(a, dtype=np.float64)
            setattr(o, name, _p(arrs[name]))
        self.images = images or []
        if self.images:
            ndc = np.array([len(im.gpos) for im in self.images], np.int32)
            npj = np.array([im.chi.shape[1] for im in self.images], np.int32)
            goff = np.concatenate([[0], np.cumsum(ndc)[:-1]]).astype(np.int32)
            coff = np.concatenate([[0], np.cumsum(ndc.astype(np.int64) * npj)[:-1]]).astype(np.int32)
            arrs["img_atom"] = np.array([im.atom for im in self.images], np.int32)
            arrs["img_ndc"] = ndc
            arrs["img_gpos_off"] = goff
            arrs["gpos"] = np.concatenate([im.gpos for im in self.images]).astype(np.int32)
            arrs["img_chi_off"] = coff
            arrs["img_nproj"] = npj
            arrs["chi"] = np.concatenate([np.asfortranarray(im.chi).ravel(order="F") for im in self.images])
            arrs["img_shift"] = np.concatenate([im.shift for im in self.images]).astype(np.float64)
            arrs["ip_displ"] = np.ascontiguousarray(ip_displ, np.int32)
            arrs["gamma"] = np.ascontiguousarray(gamma, np.float64)
            for k in ("img_atom", "img_ndc", "img_gpos_off", "gpos", "img_chi_off", "img_nproj",
                      "chi", "img_shift", "ip_displ", "gamma"):
                setattr(o, k, _p(arrs[k]))
            o.n_img = len(self.images)
            o.n_atom = n_atom
        o.dV = grid.dV
        for a in range(3):
            o.kpt[a] = kpt[a]
            o.cell[a] = grid.lattice.lengths[a]
        self.arrs = arrs
        self.o = o
        self.kpt = tuple(kpt)

    @property
    def is_complex(self):
        return any(abs(k) > 0 for k in self.kpt)

    def set_kpoint(self, kpt):
        for a in range(3):
            self.o.kpt[a] = kpt[a]
        self.kpt = tuple(kpt)

    # -- grid-sized loops --------------------------------------------------
    def halo(self, x, complex_=False):
        nd = self.grid.nd
        x = np.ascontiguousarray(x)
        ncol = x.size // nd
        R = self.st.fdn
        vol = (self.grid.nx + 2 * R) * (self.grid.ny + 2 * R) * (self.grid.nz + 2 * R)
        if complex_:
            x = x.astype(np.complex128)
            out = np.zeros(vol * ncol, np.complex128)
            lib().orc_halo_cplx(ctypes.byref(self.o), _p(x), _p(out), ncol)
        else:
            x = x.astype(np.float64)
            out = np.zeros(vol * ncol, np.float64)
            lib().orc_halo_real(ctypes.byref(self.o), _p(x), _p(out), ncol)
        return out

    def apply(self, psi, veff, c=0.0, nonlocal_=True, complex_=None):
        """Hamiltonian.cpp:64-106 on column-major (nd, ncol) data."""
        psi = np.asfortranarray(psi)
        if complex_ is None:
            complex_ = np.iscomplexobj(psi) or self.is_complex
        nd = self.grid.nd
        ncol = psi.size // nd
        veff = np.ascontiguousarray(veff, np.float64) if veff is not None else None
        if complex_:
            x = np.asfortranarray(psi, np.complex128)
            y = np.zeros_like(x)
            lib().orc_h_apply_cplx(ctypes.byref(self.o), _p(x), _p(veff), _p(y), ncol,
                                   ctypes.c_double(c), int(nonlocal_ and bool(self.images)))
        else:
            x = np.asfortranarray(psi, np.float64)
            y = np.zeros_like(x)
            lib().orc_h_apply_real(ctypes.byref(self.o), _p(x), _p(veff), _p(y), ncol,
                                   ctypes.c_double(c), int(nonlocal_ and bool(self.images)))
        return y

    def cheb_filter(self, x, veff, lambda_c, eig_min, eig_max, degree, nonlocal_=True, threads=1):
        """EigenSolver.cpp:558-619.  threads > 1 (real blocks) shares the
        columns among OpenMP threads; the result is the same bit for bit."""
        complex_ = np.iscomplexobj(x) or self.is_complex
        nd = self.grid.nd
        ncol = x.size // nd
        veff = np.ascontiguousarray(veff, np.float64)
        dbl = ctypes.c_double
        if complex_:
            xx = np.asfortranarray(x, np.complex128)
            y = np.zeros_like(xx)
            lib().orc_cheb_filter_cplx(ctypes.byref(self.o), _p(xx), _p(y), _p(veff), ncol,
                                       dbl(lambda_c), dbl(eig_min), dbl(eig_max), degree,
                                       int(nonlocal_ and bool(self.images)))
        else:
            xx = np.asfortranarray(x, np.float64)
            y = np.zeros_like(xx)
            if threads > 1:
                lib().orc_cheb_filter_real_par(ctypes.byref(self.o), _p(xx), _p(y), _p(veff), ncol,
                                               dbl(lambda_c), dbl(eig_min), dbl(eig_max), degree,
                                               int(nonlocal_ and bool(self.images)), int(threads))
            else:
                lib().orc_cheb_filter_real(ctypes.byref(self.o), _p(xx), _p(y), _p(veff), ncol,
                                           dbl(lambda_c), dbl(eig_min), dbl(eig_max), degree,
                                           int(nonlocal_ and bool(self.images)))
        return y

    def lanczos_bounds(self, veff, tol=1e-2, max_iter=1000, seed=1, nonlocal_=True, complex_=None):
        """Lanczos.hpp:104-198 (returns eig_min-0.1, eig_max*1.01, iterations).
        complex_: the complex (k-point) form, also at k = 0; default: a
        non-zero k-point."""
        veff = np.ascontiguousarray(veff, np.float64)
        mn = ctypes.c_double(); mx = ctypes.c_double(); it = ctypes.c_int()
        cplx = self.is_complex if complex_ is None else complex_
        fn = lib().orc_lanczos_cplx if cplx else lib().orc_lanczos_real
        fn(ctypes.byref(self.o), _p(veff), ctypes.c_double(tol), max_iter, ctypes.c_uint(seed),
           int(nonlocal_ and bool(self.images)), ctypes.byref(mn), ctypes.byref(mx), ctypes.byref(it))
        return mn.value, mx.value, it.value

    # -- dense subspace steps (EigenSolver.cpp:621-681) -----------------------
    def solve(self, psi, veff, lambda_c, eig_min, eig_max, degree=20, nonlocal_=True):
        """EigenSolver.cpp:738-818: filter, Cholesky-QR, Rayleigh-Ritz, rotate.
        Returns (psi_out, eigvals, new lambda_cutoff)."""
        import scipy.linalg as sla
        dV = self.grid.dV
        Y = self.cheb_filter(psi, veff, lambda_c, eig_min, eig_max, degree, nonlocal_)
        S = dV * (Y.conj().T @ Y)
        U = sla.cholesky(S, lower=False)
        # X <- X U^-1  (trsm 'R','U','N','N'): solve U^T Z = X^T, X = Z^T
        Y = sla.solve_triangular(U, Y.T, trans="T", lower=False).T
        HY = self.apply(Y, veff, 0.0, nonlocal_)
        Hs = dV * (Y.conj().T @ HY)
        Hs = 0.5 * (Hs + Hs.conj().T)
        w, Q = np.linalg.eigh(Hs)
        out = Y @ Q
        return np.asfortranarray(out), w, w[-1] + 0.1


def rand_uniform(seed, lo, hi, n):
    """glibc srand/rand stream (Wavefunction.cpp:53-62)."""
    out = np.empty(n, np.float64)
    lib().orc_rand_uniform(ctypes.c_uint(seed), ctypes.c_double(lo), ctypes.c_double(hi), _p(out),
                           ctypes.c_long(n))
    return out


def lanczos_start(complex_, n, seed=1):
    """Lanczos.hpp random_init after srand(seed), before normalisation:
    n doubles, or n complex entries whose imaginary part is drawn first."""
    out = np.empty(2 * n if complex_ else n, np.float64)
    lib().orc_lanczos_start(int(bool(complex_)), ctypes.c_long(n), ctypes.c_uint(seed), _p(out))
    return out[0::2] + 1j * out[1::2] if complex_ else out


def setup_system(grid: Grid, psps, types, frac, order=12, kpt=(0.0, 0.0, 0.0)):
    """Build the oracle operator for atoms at fractional coordinates."""
    pos = positions_noncart(grid.lattice, frac)
    images = nloc_influence(grid, psps, types, pos)
    build_chi(grid, psps, images, pos)
    ip, gam = atom_gamma(psps, types)
    return Operator(grid, order, images, len(types), ip, gam, kpt)


# ---------------------------------------------------------------------------
# SCF eigen stage (SCF.cpp:310-467, 883-935; Occupation.cpp:11-176)

def smearing(x, beta, kind=0):
    """Occupation.cpp:11-35: Fermi-Dirac (0) or 0.5 erfc (1, Gaussian).

    erfc is the C library's (Python's math.erfc calls it): in a gapped
    spectrum the Fermi-level root is flat and the last bits of the tails
    decide where the root finder stops, so the oracle uses the same erfc
    as the compiled code it is checked against."""
    import math
    x = np.asarray(x, float)
    if kind == 1:
        return 0.5 * np.vectorize(math.erfc, otypes=[float])(beta * x)
    a = x * beta
    return np.where(a > 50.0, 0.0, np.where(a < -50.0, 1.0, 1.0 / (1.0 + np.exp(np.clip(a, -50, 50)))))


def brent(f, a, b, tol=1e-14, max_iter=200):
    """Brent's root finder as NumericalMethods.cpp:14-78 takes its steps."""
    fa, fb = f(a), f(b)
    c, fc, d = a, fa, b - a
    e = d
    eps = 2.220446049250313e-16
    for _ in range(max_iter):
        if (fb > 0 and fc > 0) or (fb < 0 and fc < 0):
            c, fc = a, fa
            d = e = b - a
        if abs(fc) < abs(fb):
            a, b, c = b, c, b
            fa, fb, fc = fb, fc, fb
        tol1 = 2.0 * eps * abs(b) + 0.5 * tol
        xm = 0.5 * (c - b)
        if abs(xm) <= tol1 or fb == 0.0:
            return b
        if abs(e) >= tol1 and abs(fa) > abs(fb):
            s = fb / fa
            if a == c:
                p, q = 2.0 * xm * s, 1.0 - s
            else:
                qq, r = fa / fc, fb / fc
                p = s * (2.0 * xm * qq * (qq - r) - (b - a) * (r - 1.0))
                q = (qq - 1.0) * (r - 1.0) * (s - 1.0)
            if p > 0:
                q = -q
            p = abs(p)
            if 2.0 * p < min(3.0 * xm * q - abs(tol1 * q), abs(e * q)):
                e, d = d, p / q
            else:
                d = xm; e = d
        else:
            d = xm; e = d
        a, fa = b, fb
        b += d if abs(d) > tol1 else (tol1 if xm >= 0 else -tol1)
        fb = f(b)
    return b


def fermi_level(eigs, weights, nelectron, beta, kind=0):
    """Occupation.cpp:53-84: root of sum w f(e - Ef) = Ne on the bracket
    [emin - 10/beta, emax + 10/beta] (else 100/beta), Brent to 1e-14."""
    e = np.asarray(eigs, float)
    w = np.asarray(weights, float)
    # summed in order, as Occupation.cpp:36-46 does: near a flat root the
    # rounding of the sum is part of where the root finder stops
    g = lambda ef: float(np.cumsum(w * smearing(e - ef, beta, kind))[-1] - nelectron)  # noqa: E731
    a, b = e.min() - 10.0 / beta, e.max() + 10.0 / beta
    if g(a) * g(b) > 0:
        a, b = e.min() - 100.0 / beta, e.max() + 100.0 / beta
    return brent(g, a, b)


class ScfEigStage:
    """SCF::estimate_spectral_bounds (SCF.cpp:883-935) + SCF::solve_eigenproblem
    (SCF.cpp:310-467) for nspin spins and the k-points of one operator, FP64
    throughout.  veff holds nspin x Nd values (spin-major); psis holds the
    nspin x nkpt blocks in the order s * nkpt + k."""

    def __init__(self, op, nband, nelectron, beta, kpts=None, kpt_weights=None, degree=20, rho_trigger=4,
                 nchefsi=1, kind=0, nspin=1):
        self.op, self.nband, self.ne, self.beta, self.kind = op, nband, nelectron, beta, kind
        self.kpts = [None] if kpts is None else list(kpts)
        nk = len(self.kpts)
        self.kw = np.full(nk, 1.0 / nk) if kpt_weights is None else np.asarray(kpt_weights, float)
        self.degree, self.rho_trigger, self.nchefsi = degree, rho_trigger, nchefsi
        self.nspin = nspin
        self.bounds = None
        self.solves = 0

    def step(self, scf_iter, veff, psis):
        op = self.op
        nd = op.grid.nd
        veff = np.asarray(veff, float).reshape(self.nspin, nd)
        nk = len(self.kpts)
        if self.bounds is None or scf_iter == 0:
            # SCF.cpp:912-929: Lanczos per spin at the first k-point, once
            # per SCF run (at its first iteration)
            self.eig_min = np.zeros(self.nspin)
            self.eig_max = np.zeros(self.nspin)
            for s in range(self.nspin):
                if self.kpts[0] is not None:
                    op.set_kpoint(self.kpts[0])
                self.eig_min[s], self.eig_max[s], _ = op.lanczos_bounds(veff[s],
                                                                        complex_=self.kpts[0] is not None)
            self.lambda_c = 0.5 * (self.eig_min[0] + self.eig_max[0])
            self.bounds = True
        # SCF.cpp:323: rho_trigger passes in the first iteration
        passes = self.rho_trigger if scf_iter == 0 else self.nchefsi
        eigs = [None] * (self.nspin * nk)
        for _ in range(passes):
            for s in range(self.nspin):
                for k, kp in enumerate(self.kpts):
                    if kp is not None:
                        op.set_kpoint(kp)
                    i = s * nk + k
                    psis[i], eigs[i], _ = op.solve(psis[i], veff[s], self.lambda_c, self.eig_min[s],
                                                   self.eig_max[s], self.degree)
                    self.solves += 1
            # SCF.cpp:436-449: eig_min per spin from k-point 0 (iterations
            # after the first), lambda_c from the largest last eigenvalue
            last_max = -1e30
            for s in range(self.nspin):
                e0 = eigs[s * nk]
                if scf_iter > 0:
                    self.eig_min[s] = e0[0]
                last_max = max(last_max, e0[-1])
            self.lambda_c = last_max + 0.1
            # Occupation.cpp:86-176: spin factor 2 for one spin
            spin_fac = 2.0 if self.nspin == 1 else 1.0
            e = np.concatenate(eigs)
            w = np.concatenate([np.full(self.nband, spin_fac * self.kw[k]) for s in range(self.nspin)
                                for k in range(nk)])
            self.fermi = fermi_level(e, w, self.ne, self.beta, self.kind)
            self.occ = smearing(e - self.fermi, self.beta, self.kind)
            self.weights = w
        return psis, np.concatenate(eigs)
