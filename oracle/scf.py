"""CPU oracle of the SCF tail around the CheFSI path -- TEST INFRASTRUCTURE ONLY.

Imported only by ``tests/`` and ``bench.py``'s CPU-baseline leg, as the
checker; the product library never loads it.  numpy restatements of:

  * the local-pseudopotential set-up: atom boxes (Crystal.cpp:45-134),
    pseudocharge, reference charge, Vc, Eself, Ec, atomic and core densities
    (Electrostatics.cpp:16-520);
  * exchange-correlation: Slater exchange, Perdew-Zunger 1981 and
    Perdew-Wang 1992 correlation, PBE (Perdew, Burke & Ernzerhof, PRL 77,
    3865 (1996)), written from the published formulas (the reference calls
    libxc, XCFunctional.cpp:44-340; libxc is absent here);
  * the AAR Poisson solve (PoissonSolver.cpp:43-166) and the Kerker
    preconditioner (KerkerPreconditioner.cpp:37-95, LinearSolver.cpp:23-130),
    and the same two equations solved exactly in the Fourier basis of the
    periodic grid (fd_symbol; the library's LYNX_SOLVE_SPECTRAL);
  * Pulay mixing with the density constraint (Mixer.cpp:116-266);
  * density, energies and the SCF loop (ElectronDensity.cpp:55-140,
    Energy.cpp:140-200, SCF.cpp:110-291, 470-790);
  * atomic forces: local (pseudocharge), nonlocal (Kleinman-Bylander
    projectors) and the core-correction XC term (Forces.cpp:82-420,
    509-800, 1108-1288).
"""
from __future__ import annotations

import math

import numpy as np

from . import Grid, Lattice, Operator, positions_noncart, stencil_coeffs  # noqa: F401


# ---------------------------------------------------------------------------
# local pseudopotential set-up

RC_REF = 0.5   # cutoff of the smooth reference potential (Electrostatics.cpp:176)

# What fixes each function here independently of its own formula
# (tests/test_oracle_scf.py, tests/test_ref_subset.py, tests/test_gpu_scf.py).
PINNING = {
    "Electrostatics (v_ref, v_local, atom_boxes, pseudocharge, b_ref, Vc, Eself, Ec, rho_at, rho_core)":
        "the reference's own Crystal + Electrostatics compiled from source, on an orthogonal Si8 cell, the FCC "
        "Si2 cell and two-species BeCu, to 1e-12; integral of b = -Ne",
    "fx_slater": "the constants C2, C3 of tests/unit/test_XCFunctional.cpp:46-70; its derivative by finite "
                 "differences",
    "ec_pz": "the two PZ81 branches join at rs = 1 (they were fitted to); within a few mHa of PW92",
    "ec_unif (PW92)": "the published PW92 correlation energies at rs = 1, 2, 5, 10",
    "fx_pbe, fc_pbe": "zero gradient gives the LDA with PW92 (test_XCFunctional.cpp:128-150); the published "
                      "small-gradient coefficients mu = 0.21951 and beta = 0.066725 (phi^3 scaling at full "
                      "polarisation); the large-gradient limits 1 + kappa and 0",
    "xc_unpol, xc_pol": "complex-step derivatives against finite differences; zeta = 0 equals the unpolarized "
                        "forms; spin scaling of exchange",
    "aar": "the closed-form FD Poisson solution of a plane wave",
    "fd_symbol": "the FFT of the stencil's own -lap x equals the symbol times the FFT of x (random x, uneven "
                 "orthogonal grid and FCC cell)",
    "ScfOracle (spectral vs aar)": "both converge to the same SCF energy (1e-8 Ha/atom)",
    "ScfOracle.kerker_apply": "z solves (-lap + kTF^2) z = -amix (lap - idiemac kTF^2) r",
    "ScfOracle.mix": "the density constraint: non-negative, integral Ne",
    "ScfOracle.run": "converges; the library following it iteration by iteration reproduces the reference's "
                     "published BeCu energy (test_suite/systems/BeCu_bulk/reference.json) to 1e-7 Ha/atom",
    "ScfOracle.initial_density / run(rho0, mag0)": "the atomic density passed as a restart reproduces the default "
                     "run's history bit for bit; a restart from the converged density stops at the minimum "
                     "iteration count on the same energy; the zero floor of (total +- mag) / 2 checked against "
                     "hand values",
}


def v_ref(r, rc, Z):
    """Electrostatics.cpp:16-33: -Z/r outside rc, a polynomial inside."""
    r = np.asarray(r, float)
    rc2 = rc * rc; rc3 = rc2 * rc; rc6 = rc3 * rc3; rc7 = rc6 * rc; rc8 = rc7 * rc
    a0, a3, a5, a6, a7 = 2.4 / rc, -2.8 / rc3, 5.6 / rc6, -6.0 / rc7, 1.8 / rc8
    r2 = r * r; r3 = r2 * r
    inner = -Z * (r2 * (r3 * (a5 + r * (a6 + r * a7)) + a3) + a0)
    with np.errstate(divide="ignore"):
        outer = -Z / np.where(r > rc, r, 1.0)
    return np.where(r > rc, outer, inner)


def _spline1(x, y, yd, r):
    """NumericalMethods.cpp:194-225 on an array of points."""
    from . import spline_eval
    return spline_eval(x, y, yd, r)


def v_local(ps, r):
    """Electrostatics.cpp:216-223: Vloc(0) at the nucleus, spline(r Vloc)/r on
    the table, -Z/r beyond."""
    r = np.asarray(r, float)
    out = np.empty_like(r)
    small = r < 1e-10
    inside = (~small) & (r < ps.r[-1])
    far = (~small) & ~inside
    out[small] = ps.vloc0
    out[inside] = _spline1(ps.r, ps.r_vloc, ps.r_vloc_d, r[inside]) / r[inside]
    out[far] = -ps.zion / r[far]
    return out


def atom_boxes(grid: Grid, types, pos, rc_max):
    """Crystal.cpp:45-134: (type, atom, image position, start, end) for every
    periodic image of every atom whose rc_max box meets the grid."""
    lat = grid.lattice
    n = (grid.nx, grid.ny, grid.nz)
    ext = np.ones(3) if lat.orthogonal else lat.nc_extent
    out = []
    for it in range(max(types) + 1):
        for ia, t in enumerate(types):
            if t != it:
                continue
            nimg = [int(math.ceil(rc_max * ext[a] / lat.lengths[a])) for a in range(3)]
            for iz in range(-nimg[2], nimg[2] + 1):
                for iy in range(-nimg[1], nimg[1] + 1):
                    for ix in range(-nimg[0], nimg[0] + 1):
                        c = pos[ia] + np.array([ix, iy, iz]) * lat.lengths
                        s, e = [], []
                        for a in range(3):
                            if lat.orthogonal:
                                sa = math.ceil((c[a] - rc_max) / grid.h[a])
                                ea = math.floor((c[a] + rc_max) / grid.h[a])
                            else:
                                rr = rc_max * ext[a] / grid.h[a]
                                sa = math.ceil(c[a] / grid.h[a] - rr)
                                ea = math.floor(c[a] / grid.h[a] + rr)
                            s.append(max(sa, 0)); e.append(min(ea, n[a] - 1))
                        if all(s[a] <= e[a] for a in range(3)):
                            out.append((it, ia, c, s, e))
    return out


def _dist(lat, rx, ry, rz):
    if lat.orthogonal:
        return np.sqrt(rx * rx + ry * ry + rz * rz)
    d = np.stack([rx, ry, rz], axis=-1) @ lat.uvec
    return np.sqrt((d * d).sum(axis=-1))


def _lap_box(V, st, R, orth, coef):
    """Electrostatics.cpp:48-127: coef * FD Laplacian of V (extended box,
    indexed [k, j, i]) on its inner region, the mixed terms unfactored."""
    nzp, nyp, nxp = V.shape
    I = slice(R, nxp - R); J = slice(R, nyp - R); K = slice(R, nzp - R)

    def sh(dk, dj, di):
        return V[R + dk:nzp - R + dk, R + dj:nyp - R + dj, R + di:nxp - R + di]

    cx, cy, cz = st.D2
    val = (cx[0] + cy[0] + cz[0]) * coef * V[K, J, I]
    for p in range(1, R + 1):
        val = val + coef * cx[p] * (sh(0, 0, p) + sh(0, 0, -p))
        val = val + coef * cy[p] * (sh(0, p, 0) + sh(0, -p, 0))
        val = val + coef * cz[p] * (sh(p, 0, 0) + sh(-p, 0, 0))
    if not orth:
        cxy, cxz, cyz = st.D2m
        d1y, d1z = st.D1[1], st.D1[2]
        for p in range(1, R + 1):
            for q in range(1, R + 1):
                val = val + coef * cxy[p] * d1y[q] * (sh(0, q, p) - sh(0, -q, p) - sh(0, q, -p) + sh(0, -q, -p))
        for p in range(1, R + 1):
            for q in range(1, R + 1):
                val = val + coef * cxz[p] * d1z[q] * (sh(q, 0, p) - sh(-q, 0, p) - sh(q, 0, -p) + sh(-q, 0, -p))
        for p in range(1, R + 1):
            for q in range(1, R + 1):
                val = val + coef * cyz[p] * d1z[q] * (sh(q, p, 0) - sh(-q, p, 0) - sh(q, -p, 0) + sh(-q, -p, 0))
    return val


def _d1_box(V, st, R, a):
    """Forces.cpp:371-376: FD first derivative along lattice axis a on the
    inner region of an extended box indexed [k, j, i]."""
    nzp, nyp, nxp = V.shape
    K, J, I = slice(R, nzp - R), slice(R, nyp - R), slice(R, nxp - R)
    out = np.zeros((nzp - 2 * R, nyp - 2 * R, nxp - 2 * R))
    d1 = st.D1[a]
    for p in range(1, R + 1):
        if a == 0:
            out += (V[K, J, R + p:nxp - R + p] - V[K, J, R - p:nxp - R - p]) * d1[p]
        elif a == 1:
            out += (V[K, R + p:nyp - R + p, I] - V[K, R - p:nyp - R - p, I]) * d1[p]
        else:
            out += (V[R + p:nzp - R + p, J, I] - V[R - p:nzp - R - p, J, I]) * d1[p]
    return out


def grad_periodic(x, grid, st, a, kpt=None):
    """FD first derivative along lattice axis a of a periodic field (complex:
    Bloch field, wrapped values times e^{+-i k_a L_a}), HaloExchange.cpp
    34-189 then Gradient::apply."""
    X = np.asarray(x).reshape(grid.nz, grid.ny, grid.nx)
    ax = {0: 2, 1: 1, 2: 0}[a]
    n = X.shape[ax]
    ph = 1.0
    if kpt is not None:
        ph = np.exp(1j * kpt[a] * grid.lattice.lengths[a])
    out = np.zeros_like(X)
    idx = np.arange(n)
    for p in range(1, st.fdn + 1):
        up = np.take(X, (idx + p) % n, axis=ax)
        dn = np.take(X, (idx - p) % n, axis=ax)
        if kpt is not None:
            shp = [1, 1, 1]
            shp[ax] = n
            up = up * np.where(idx + p >= n, ph, 1.0).reshape(shp)
            dn = dn * np.where(idx - p < 0, np.conj(ph), 1.0).reshape(shp)
        out = out + st.D1[a][p] * (up - dn)
    return out.ravel()


class Electrostatics:
    """Pseudocharge, reference charge, Vc, Eself, Ec, atomic and core
    densities of the atoms at fractional coordinates (type-major order)."""

    def __init__(self, grid: Grid, psps, types, frac, order=12):
        lat = grid.lattice
        st = stencil_coeffs(order, grid)
        R = st.fdn
        pos = positions_noncart(lat, frac)
        types = list(types)
        nx, ny, nz = grid.nx, grid.ny, grid.nz
        rc_max = max(max(max(p.rc), p.r[-1]) for p in psps) + 8.0 * grid.h.max()
        shape = (nz, ny, nx)
        b = np.zeros(shape); bref = np.zeros(shape); vc = np.zeros(shape)
        rho_at = np.zeros(shape); rho_core = np.zeros(shape)
        eself = 0.0
        coef = -0.25 / math.pi
        self.boxes = atom_boxes(grid, types, pos, rc_max)
        for it, ia, c, s, e in self.boxes:
            ps = psps[it]
            gi = np.arange(s[0] - R, e[0] + R + 1)
            gj = np.arange(s[1] - R, e[1] + R + 1)
            gk = np.arange(s[2] - R, e[2] + R + 1)
            rz, ry, rx = np.meshgrid(gk * grid.h[2] - c[2], gj * grid.h[1] - c[1], gi * grid.h[0] - c[0],
                                     indexing="ij")
            r = _dist(lat, rx, ry, rz)
            VJ = v_local(ps, r.ravel()).reshape(r.shape)
            VR = v_ref(r, RC_REF, ps.zion)
            bj = _lap_box(VJ, st, R, lat.orthogonal, coef)
            br = _lap_box(VR, st, R, lat.orthogonal, coef)
            inner = (slice(s[2], e[2] + 1), slice(s[1], e[1] + 1), slice(s[0], e[0] + 1))
            b[inner] += bj
            bref[inner] += br
            ri = r[R:-R, R:-R, R:-R] if R > 0 else r
            eself -= float((br * VR[R:-R, R:-R, R:-R]).sum())
            vc[inner] += v_ref(ri, RC_REF, ps.zion) - v_local(ps, ri.ravel()).reshape(ri.shape)
            if ps.rho_iso is not None:
                m = ri < ps.r[-1]
                add = np.zeros(ri.shape)
                add[m] = _spline1(ps.r, ps.rho_iso, ps.rho_iso_d, ri[m])
                rho_at[inner] += add
            if ps.fchrg > 1e-10 and ps.rho_c is not None:
                m = ri < ps.r[-1]
                add = np.zeros(ri.shape)
                add[m] = _spline1(ps.r, ps.rho_c, ps.rho_c_d, ri[m])
                rho_core[inner] += add
        dV = grid.dV
        self.nelectron = sum(int(psps[t].zion) for t in types)
        self.eself = 0.5 * dV * eself
        self.b = b.ravel(); self.b_ref = bref.ravel(); self.vc = vc.ravel()
        self.int_b = float(self.b.sum() * dV)
        self.ec = 0.5 * dV * float(((self.b + self.b_ref) * self.vc).sum())
        rho_at = rho_at.ravel()
        tot = rho_at.sum() * dV
        self.rho_at = rho_at * (self.nelectron / tot) if tot > 1e-10 else rho_at
        self.rho_core = rho_core.ravel()
        self.has_nlcc = any(p.fchrg > 1e-10 for p in psps)
        self.grid, self.st, self.psps, self.types, self.pos = grid, st, psps, types, pos

    def local_forces(self, phi, vxc0=None):
        """Forces.cpp:186-420 and 1108-1275, per atom (lattice-aligned
        components): -int b_J grad(phi) + (1/2) int [grad(Vc_J)(b + b_ref)
        - grad(Vc)(b_J + b_J,ref)], and with a core density int Vxc
        grad(rho_c,J) (the first spin's Vxc for two spins, as the
        reference passes SCF::Vxc())."""
        g, st = self.grid, self.st
        lat = g.lattice
        R = st.fdn
        shape = (g.nz, g.ny, g.nx)
        dphi = [grad_periodic(phi, g, st, a).reshape(shape) for a in range(3)]
        dvc = [grad_periodic(self.vc, g, st, a).reshape(shape) for a in range(3)]
        bb = (self.b + self.b_ref).reshape(shape)
        coef = -0.25 / math.pi
        natom = len(self.types)
        f_loc = np.zeros((natom, 3))
        f_xc = np.zeros((natom, 3))
        for it, ia, c, s, e in self.boxes:
            ps = self.psps[it]
            gi = np.arange(s[0] - R, e[0] + R + 1)
            gj = np.arange(s[1] - R, e[1] + R + 1)
            gk = np.arange(s[2] - R, e[2] + R + 1)
            rz, ry, rx = np.meshgrid(gk * g.h[2] - c[2], gj * g.h[1] - c[1], gi * g.h[0] - c[0], indexing="ij")
            r = _dist(lat, rx, ry, rz)
            VJ = v_local(ps, r.ravel()).reshape(r.shape)
            VR = v_ref(r, RC_REF, ps.zion)
            bj = _lap_box(VJ, st, R, lat.orthogonal, coef)
            br = _lap_box(VR, st, R, lat.orthogonal, coef)
            inner = (slice(s[2], e[2] + 1), slice(s[1], e[1] + 1), slice(s[0], e[0] + 1))
            for a in range(3):
                dvcj = _d1_box(VR - VJ, st, R, a)
                f = -float((bj * dphi[a][inner]).sum())
                corr = float((dvcj * bb[inner]).sum() - (dvc[a][inner] * (bj + br)).sum())
                f_loc[ia, a] += (f + 0.5 * corr) * g.dV
            if vxc0 is not None and ps.fchrg > 1e-10 and ps.rho_c is not None:
                rc = np.zeros(r.shape)
                m = r <= ps.r[-1]
                rc[m] = _spline1(ps.r, ps.rho_c, ps.rho_c_d, r[m])
                v = np.asarray(vxc0).reshape(shape)[inner]
                for a in range(3):
                    f_xc[ia, a] += float((v * _d1_box(rc, st, R, a)).sum()) * g.dV
        return f_loc, f_xc


# ---------------------------------------------------------------------------
# exchange-correlation (energy densities rho * eps; derivatives by the
# complex step, an independent differentiation of the same energies)

_PI = math.pi


def _cbrt(x):
    return x ** (1.0 / 3.0) if np.iscomplexobj(x) else np.cbrt(x)


def fx_slater(rho):
    """Slater exchange, unpolarized: rho eps_x = -(3/4)(3/pi)^(1/3) rho^(4/3)."""
    return -0.75 * (3.0 / _PI) ** (1.0 / 3.0) * rho * _cbrt(rho)


def ec_pz(rs, pol):
    """Perdew & Zunger, PRB 23, 5048 (1981), Table XII."""
    g, b1, b2 = (-0.0843, 1.3981, 0.2611) if pol else (-0.1423, 1.0529, 0.3334)
    A, B, C, D = (0.01555, -0.0269, 0.0007, -0.0048) if pol else (0.0311, -0.048, 0.0020, -0.0116)
    hi = g / (1.0 + b1 * np.sqrt(rs) + b2 * rs)
    lr = np.log(rs)
    lo = A * lr + B + C * rs * lr + D * rs
    return np.where(np.real(rs) >= 1.0, hi, lo)


PW92 = {False: ((0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294),
                (0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517),
                (0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671), 1.709921),
        True: ((0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294),
               (0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517),
               (0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671), 1.709920934161365617563962776245)}


def pw_g(rs, par):
    """Perdew & Wang, PRB 45, 13244 (1992), eq. 10 with p = 1."""
    A, a1, b1, b2, b3, b4 = par
    s = np.sqrt(rs)
    return -2.0 * A * (1.0 + a1 * rs) * np.log(1.0 + 1.0 / (2.0 * A * (b1 * s + b2 * rs + b3 * rs * s + b4 * rs * rs)))


def fzeta(z):
    return ((1 + z) * _cbrt(1 + z) + (1 - z) * _cbrt(1 - z) - 2.0) / (2.0 * 2.0 ** (1.0 / 3.0) - 2.0)


def ec_unif(rs, zeta, kind, mod=False):
    """Correlation per particle: PZ81 (kind 0) or PW92 (kind 1; mod: the
    more precise A values PBE uses); zeta None for unpolarized."""
    if kind == 0:
        eu = ec_pz(rs, False)
        return eu if zeta is None else eu + fzeta(zeta) * (ec_pz(rs, True) - eu)
    p0, p1, p2, fz20 = PW92[mod]
    e0 = pw_g(rs, p0)
    if zeta is None:
        return e0
    e1 = pw_g(rs, p1)
    ac = -pw_g(rs, p2)
    f = fzeta(zeta)
    z4 = zeta ** 4
    return e0 + ac * f * (1 - z4) / fz20 + (e1 - e0) * f * z4


KAPPA, BETA_PBE = 0.8040, 0.06672455060314922
MU_PBE = BETA_PBE * _PI * _PI / 3.0
GAMMA_PBE = (1.0 - math.log(2.0)) / (_PI * _PI)


def fx_pbe(rho, sigma):
    """PBE exchange of an unpolarized density, PRL 77, 3865 (1996) eqs. 13-14."""
    kf = _cbrt(3.0 * _PI * _PI * rho)
    s2 = sigma / (4.0 * kf * kf * rho * rho)
    return fx_slater(rho) * (1.0 + KAPPA - KAPPA / (1.0 + MU_PBE * s2 / KAPPA))


def fc_pbe(rho, sigma, zeta=None):
    """PBE correlation, eqs. 3, 7, 8, with PW92 (precise A values) for the
    uniform gas."""
    rs = _cbrt(3.0 / (4.0 * _PI * rho))
    ec = ec_unif(rs, zeta, 1, mod=True)
    phi = 1.0 if zeta is None else 0.5 * (_cbrt(1 + zeta) ** 2 + _cbrt(1 - zeta) ** 2)
    kf = _cbrt(3.0 * _PI * _PI * rho)
    ks2 = 4.0 * kf / _PI
    t2 = sigma / (4.0 * phi * phi * ks2 * rho * rho)
    A = (BETA_PBE / GAMMA_PBE) / (np.exp(-ec / (GAMMA_PBE * phi ** 3)) - 1.0)
    At2 = A * t2
    H = GAMMA_PBE * phi ** 3 * np.log(1.0 + (BETA_PBE / GAMMA_PBE) * t2 * (1 + At2) / (1 + At2 + At2 * At2))
    return rho * (ec + H)


def _f_unpol(kind, rho, sigma):
    if kind == 2:
        return fx_pbe(rho, sigma) + fc_pbe(rho, sigma)
    return fx_slater(rho) + rho * ec_unif(_cbrt(3.0 / (4.0 * _PI * rho)), None, kind)


def _f_pol(kind, ru, rd, suu, sud, sdd):
    r = ru + rd
    z = (ru - rd) / r
    if kind == 2:
        fx = 0.5 * fx_pbe(2 * ru, 4 * suu) + 0.5 * fx_pbe(2 * rd, 4 * sdd)
        return fx + fc_pbe(r, suu + 2 * sud + sdd, z)
    return 0.5 * fx_slater(2 * ru) + 0.5 * fx_slater(2 * rd) + r * ec_unif(_cbrt(3.0 / (4.0 * _PI * r)), z, kind)


def _cstep(f, args, k):
    """d f / d args[k] by the complex step."""
    h = 1e-20
    a = [np.asarray(x, complex) for x in args]
    scale = np.maximum(1.0, np.abs(a[k].real))
    a[k] = a[k] + 1j * h * scale
    return f(*a).imag / (h * scale)


RHO_MIN = 1e-15


def xc_unpol(kind, rho, sigma=None):
    """exc (per particle), vrho, v2 = 2 d(rho eps)/d sigma (0 for LDA)."""
    rho = np.asarray(rho, float)
    sig = np.zeros_like(rho) if sigma is None else np.maximum(np.asarray(sigma, float), 1e-40)
    ok = rho > RHO_MIN
    r, s = rho[ok], sig[ok]
    f = lambda a, b: _f_unpol(kind, a, b)  # noqa: E731
    exc = np.zeros_like(rho); vr = np.zeros_like(rho); v2 = np.zeros_like(rho)
    exc[ok] = f(r, s).real / r
    vr[ok] = _cstep(f, [r, s], 0)
    if kind == 2:
        v2[ok] = 2.0 * _cstep(f, [r, s], 1)
    return exc, vr, v2


def xc_pol(kind, ru, rd, suu=None, sud=None, sdd=None):
    """exc (per particle of the total), v_up, v_dn and (PBE) vsigma_uu,
    vsigma_ud, vsigma_dd."""
    ru = np.asarray(ru, float); rd = np.asarray(rd, float)
    z = np.zeros_like(ru)
    suu = z if suu is None else np.maximum(np.asarray(suu, float), 1e-40)
    sdd = z if sdd is None else np.maximum(np.asarray(sdd, float), 1e-40)
    sud = z if sud is None else np.asarray(sud, float)
    ok = ru + rd > RHO_MIN
    args = [x[ok] for x in (ru, rd, suu, sud, sdd)]
    f = lambda *a: _f_pol(kind, *a)  # noqa: E731
    out = [np.zeros_like(ru) for _ in range(6)]
    out[0][ok] = f(*args).real / (args[0] + args[1])
    for k in range(5 if kind == 2 else 2):
        out[1 + k][ok] = _cstep(f, args, k)
    return tuple(out)


# ---------------------------------------------------------------------------
# linear solves, potential, mixing, density, energy, SCF loop

AAR_M, AAR_P, AAR_OMEGA, AAR_BETA = 7, 6, 0.6, 0.6


def gauss_solve(A, b):
    """NumericalMethods.cpp:84-120 (partial pivoting, vanishing pivots skipped)."""
    A = np.array(A, float); b = np.array(b, float)
    n = len(b)
    for k in range(n):
        p = k + int(np.argmax(np.abs(A[k:, k])))
        if p != k:
            A[[k, p]] = A[[p, k]]; b[[k, p]] = b[[p, k]]
        d = A[k, k]
        if abs(d) < 1e-14:
            continue
        for i in range(k + 1, n):
            f = A[i, k] / d
            A[i, k + 1:] -= f * A[k, k + 1:]
            b[i] -= f * b[k]
    for k in range(n - 1, -1, -1):
        if abs(A[k, k]) < 1e-14:
            continue
        b[k] -= A[k, k + 1:] @ b[k + 1:]
        b[k] /= A[k, k]
    return b


def aar(apply_A, b, x, tol, m_inv, max_iter):
    """LinearSolver.cpp:23-130 / PoissonSolver.cpp:43-130: Richardson steps
    with weight omega, an Anderson step every p-th iteration over the last
    m differences, the residual norm checked only after Anderson steps."""
    x = np.array(x, float)
    n = len(b)
    X = np.zeros((AAR_M, n)); F = np.zeros((AAR_M, n))
    xo = x.copy(); fo = np.zeros(n)
    abs_tol = tol * math.sqrt(float(b @ b))
    r = b - apply_A(x)
    rnorm = abs_tol + 1.0
    it = 0
    while rnorm > abs_tol and it < max_iter:
        f = m_inv * r
        if it > 0:
            h = (it - 1) % AAR_M
            X[h] = x - xo; F[h] = f - fo
        xo = x.copy(); fo = f.copy()
        if (it + 1) % AAR_P == 0 and it > 0:
            cols = min(it, AAR_M)
            FtF = F[:cols] @ F[:cols].T
            g = gauss_solve(FtF, F[:cols] @ f)
            x = xo + AAR_BETA * f - g @ (X[:cols] + AAR_BETA * F[:cols])
            r = b - apply_A(x)
            rnorm = math.sqrt(float(r @ r))
        else:
            x = xo + AAR_OMEGA * f
            r = b - apply_A(x)
        it += 1
    return x, it


def fd_symbol(grid, st):
    """Fourier symbol of the periodic FD operator -lap (the stencil of
    Hamiltonian.cpp:126-267 with its mixed terms) on numpy's rfftn layout
    (nz, ny, nx // 2 + 1): a plane wave with phases theta_a per grid step
    picks up 2 cos(p theta) from each symmetric pair and (2i sin)(2i sin)
    from each mixed four-point group."""
    nx, ny, nz, R = grid.nx, grid.ny, grid.nz, st.fdn
    tx = 2 * np.pi * np.arange(nx // 2 + 1) / nx
    ty = 2 * np.pi * np.arange(ny) / ny
    tz = 2 * np.pi * np.arange(nz) / nz
    p = np.arange(1, R + 1)

    def cos_sum(c, t):
        return c[0] + 2.0 * (np.cos(np.outer(t, p)) * np.asarray(c)[1:]).sum(axis=1)

    def sin_sum(c, t):
        return (np.sin(np.outer(t, p)) * np.asarray(c)[1:]).sum(axis=1)

    lap = (cos_sum(st.D2[0], tx)[None, None, :] + cos_sum(st.D2[1], ty)[None, :, None]
           + cos_sum(st.D2[2], tz)[:, None, None])
    lap = lap - 4.0 * (sin_sum(st.D2m[0], tx)[None, None, :] * sin_sum(st.D1[1], ty)[None, :, None]
                       + sin_sum(st.D2m[1], tx)[None, None, :] * sin_sum(st.D1[2], tz)[:, None, None]
                       + sin_sum(st.D2m[2], ty)[None, :, None] * sin_sum(st.D1[2], tz)[:, None, None])
    return -lap


class ScfOracle:
    """SCF::run (SCF.cpp:110-291) in numpy around the oracle eigen stage:
    density mixing with Pulay + Kerker, collinear spin, LDA (PZ, PW) or PBE.
    solver "aar" solves the Poisson and Kerker equations as the reference
    does; "spectral" solves them exactly with fd_symbol."""

    KTF, IDIEMAC = 1.0, 0.1

    def __init__(self, grid, psps, types, frac, nband, beta, xc=2, nspin=1, kpts=None, kpt_weights=None,
                 smearing=0, degree=20, rho_trigger=4, nchefsi=1, tol=1e-6, max_iter=100, min_iter=2,
                 mixing_history=7, mixing_param=0.3, kerker=True, order=12, solver="aar"):
        from . import ScfEigStage, setup_system
        self.grid = grid
        self.st = stencil_coeffs(order, grid)
        self.op = setup_system(grid, psps, list(types), frac, order)
        self.lap_op = Operator(grid, order)
        self.E = Electrostatics(grid, psps, types, frac, order)
        self.nband, self.beta, self.xc, self.nspin = nband, beta, xc, nspin
        self.kpts = [None] if kpts is None else [tuple(k) for k in kpts]
        nk = len(self.kpts)
        self.kw = np.full(nk, 1.0 / nk) if kpt_weights is None else np.asarray(kpt_weights, float)
        self.smearing, self.tol, self.max_iter, self.min_iter = smearing, tol, max_iter, min_iter
        self.m, self.amix, self.kerker = mixing_history, mixing_param, kerker
        self.stage = ScfEigStage(self.op, nband, self.E.nelectron, beta, None if kpts is None else self.kpts,
                                 self.kw, degree, rho_trigger, nchefsi, smearing, nspin)
        h = grid.h
        heff = h[0] if (abs(h[0] - h[1]) < 1e-12 and abs(h[1] - h[2]) < 1e-12) else \
            math.sqrt(3.0 / sum(1.0 / (x * x) for x in h))
        self.poisson_tol = tol * 0.01
        self.precond_tol = heff * heff * 1e-3
        self.phi = np.zeros(grid.nd)
        self.diag = float(self.st.D2[0][0] + self.st.D2[1][0] + self.st.D2[2][0])
        lat = grid.lattice
        self.M = np.eye(3) if lat.orthogonal else lat.lapc_T
        self.solver = solver
        self.sym = fd_symbol(grid, self.st) if solver == "spectral" else None

    def _spectral(self, b, factor):
        g = self.grid
        B = np.fft.rfftn(b.reshape(g.nz, g.ny, g.nx))
        return np.fft.irfftn(B * factor, s=(g.nz, g.ny, g.nx), axes=(0, 1, 2)).ravel()

    # -- operators ---------------------------------------------------------
    def _A(self, x, k2):
        """-lap x + k2 x through the oracle's stencil (Veff 0, no projectors)."""
        z = np.zeros(self.grid.nd)
        return 2.0 * self.lap_op.apply(x.reshape(-1, 1), z, c=0.5 * k2, nonlocal_=False).ravel()

    def _grad(self, x, a):
        g = self.grid
        X = x.reshape(g.nz, g.ny, g.nx)
        ax = {0: 2, 1: 1, 2: 0}[a]
        d1 = self.st.D1[a]
        out = np.zeros_like(X)
        for p in range(1, self.st.fdn + 1):
            out += d1[p] * (np.roll(X, -p, axis=ax) - np.roll(X, p, axis=ax))
        return out.ravel()

    def poisson(self, rho_total):
        """EffectivePotential.cpp:44-72."""
        rhs = 4.0 * math.pi * (rho_total + self.E.b)
        if self.solver == "spectral":
            inv = np.zeros_like(self.sym)
            nz0 = np.ones(self.sym.shape, bool)
            nz0[0, 0, 0] = False
            inv[nz0] = 1.0 / self.sym[nz0]
            self.phi = self._spectral(rhs, inv)
            return self.phi
        rhs = rhs - rhs.sum() / len(rhs)
        m_inv = 1.0 if abs(self.diag) < 1e-14 else -1.0 / self.diag
        self.phi, self.poisson_iters = aar(lambda x: self._A(x, 0.0), rhs, self.phi, self.poisson_tol, m_inv, 3000)
        self.phi = self.phi - self.phi.sum() / len(self.phi)
        return self.phi

    def potential(self, rho):
        """EffectivePotential::compute; rho is (nspin == 1: total) or
        (total, up, down)."""
        nd = self.grid.nd
        core = self.E.rho_core if self.E.has_nlcc else None
        gga = self.xc == 2
        M = self.M
        if self.nspin == 1:
            r = rho[0] if core is None else np.maximum(rho[0] + core, 1e-14)
            if gga:
                g = [self._grad(r, a) for a in range(3)]
                Mg = [sum(M[i][j] * g[j] for j in range(3)) for i in range(3)]
                sigma = sum(Mg[i] * g[i] for i in range(3))
                exc, v, v2 = xc_unpol(2, r, sigma)
                for a in range(3):
                    v = v - self._grad(v2 * Mg[a], a)
            else:
                exc, v, _ = xc_unpol(self.xc, r)
            vxc = [v]
        else:
            ru, rd = rho[1].copy(), rho[2].copy()
            if core is not None:
                ru = ru + 0.5 * core; rd = rd + 0.5 * core
            ru = np.maximum(ru, 0.5e-14); rd = np.maximum(rd, 0.5e-14)
            if gga:
                gu = [self._grad(ru, a) for a in range(3)]
                gd = [self._grad(rd, a) for a in range(3)]
                Mgu = [sum(M[i][j] * gu[j] for j in range(3)) for i in range(3)]
                Mgd = [sum(M[i][j] * gd[j] for j in range(3)) for i in range(3)]
                suu = sum(Mgu[i] * gu[i] for i in range(3))
                sud = sum(Mgu[i] * gd[i] for i in range(3))
                sdd = sum(Mgd[i] * gd[i] for i in range(3))
                exc, vu, vd, vsuu, vsud, vsdd = xc_pol(2, ru, rd, suu, sud, sdd)
                for a in range(3):
                    vu = vu - self._grad(2 * vsuu * Mgu[a] + vsud * Mgd[a], a)
                    vd = vd - self._grad(vsud * Mgu[a] + 2 * vsdd * Mgd[a], a)
            else:
                exc, vu, vd = xc_pol(self.xc, ru, rd)[:3]
            vxc = [vu, vd]
        phi = self.poisson(rho[0])
        self.exc, self.vxc = exc, vxc
        self.veff = np.concatenate([v + phi for v in vxc])
        return self.veff

    def kerker_apply(self, r):
        """KerkerPreconditioner.cpp:37-95."""
        k2 = self.KTF ** 2
        if self.solver == "spectral":
            return self._spectral(r, self.amix * (self.sym + self.IDIEMAC * k2) / (self.sym + k2))
        Lf = -self._A(r, self.IDIEMAC * k2)         # (lap - idiemac kTF^2) r
        md = self.diag - k2
        m_inv = 1.0 if abs(md) < 1e-14 else -1.0 / md
        z, _ = aar(lambda x: self._A(x, k2), Lf, np.zeros_like(r), self.precond_tol, m_inv, 1000)
        return -self.amix * z

    def mix(self, x, g, ncol):
        """Mixer.cpp:116-266 (density mixing) with the density constraint
        (Mixer.cpp:70-114)."""
        if not hasattr(self, "_R"):
            self._it = 0
            N = len(x)
            self._fk = np.zeros(N); self._xkm1 = np.zeros(N)
            self._R = np.zeros((self.m, N)); self._F = np.zeros((self.m, N))
        fkm1 = self._fk.copy()
        fk = g - x
        if self._it > 0:
            h = (self._it - 1) % self.m
            self._R[h] = x - self._xkm1
            self._F[h] = fk - fkm1
        self._fk = fk
        if self._it > 0:
            cols = min(self._it, self.m)
            F = self._F[:cols]
            gm = gauss_solve(F @ F.T, F @ fk)
            xw = x - gm @ self._R[:cols]
            fw = fk - gm @ F
        else:
            xw, fw = x.copy(), fk.copy()
        nd = self.grid.nd
        if self.kerker:
            Pf = np.concatenate([self.kerker_apply(fw[:nd])] + [self.amix * fw[c * nd:(c + 1) * nd]
                                                                  for c in range(1, ncol)])
        else:
            Pf = self.amix * fw
        self._xkm1 = x.copy()
        x = xw + Pf
        if ncol == 2:
            up = np.maximum(0.5 * (x[:nd] + x[nd:]), 0.0)
            dn = np.maximum(0.5 * (x[:nd] - x[nd:]), 0.0)
            x = np.concatenate([up + dn, up - dn])
        else:
            x = np.maximum(x, 0.0)
        ne = x[:nd].sum() * self.grid.dV
        if ne > 1e-10:
            x = x * (self.E.nelectron / ne)
        self._it += 1
        return x

    def forces(self):
        """Forces.cpp:82-184 after run(): local + nonlocal (+ core
        correction) in lattice-aligned components, mapped to Cartesian for
        non-orthogonal cells (F = U^-1 F_nc), then the mean force removed
        (Forces.cpp:1277-1288).  Returns (total, parts) with parts the
        unsymmetrized local, nonlocal and XC terms in the same frame."""
        g, st = self.grid, self.st
        lat = g.lattice
        natom = len(self.E.types)
        vxc0 = self.vxc[0] if self.E.has_nlcc else None
        f_loc, f_xc = self.E.local_forces(self.phi, vxc0)
        f_nl = np.zeros((natom, 3))
        op = self.op
        ip = np.asarray(op.arrs["ip_displ"])
        gam = np.asarray(op.arrs["gamma"])
        tot = int(ip[-1])
        occfac = 2.0 if self.nspin == 1 else 1.0
        spn_fac = occfac * 2.0
        nk, N = len(self.kpts), self.nband
        for s in range(self.nspin):
            for k, kp in enumerate(self.kpts):
                psi = self.psis[s * nk + k]
                occ = self.occ[(s * nk + k) * N:(s * nk + k + 1) * N]
                cplx = kp is not None
                kv = np.zeros(3) if kp is None else np.asarray(kp, float)
                alpha = np.zeros((tot, N), complex if cplx else float)
                for im in op.images:
                    ph = np.exp(-1j * float(kv @ im.shift)) if cplx else 1.0
                    o = ip[im.atom]
                    alpha[o:o + im.chi.shape[1]] += ph * g.dV * (im.chi.T @ psi[im.gpos])
                for a in range(3):
                    D = np.stack([grad_periodic(psi[:, n], g, st, a, kv if cplx else None) for n in range(N)],
                                 axis=1)
                    beta = np.zeros_like(alpha)
                    for im in op.images:
                        ph = np.exp(-1j * float(kv @ im.shift)) if cplx else 1.0
                        o = ip[im.atom]
                        beta[o:o + im.chi.shape[1]] += ph * g.dV * (im.chi.T @ D[im.gpos])
                    prod = np.real(np.conj(alpha) * beta) * gam[:, None]
                    for at in range(natom):
                        fj = prod[ip[at]:ip[at + 1]].sum(axis=0)
                        f_nl[at, a] -= spn_fac * self.kw[k] * float((occ * fj).sum())
        total = f_loc + f_nl + f_xc
        parts = [f_loc.copy(), f_nl.copy(), f_xc.copy()]
        if not lat.orthogonal:
            Ui = np.linalg.inv(lat.uvec)
            total = total @ Ui.T
            parts = [p @ Ui.T for p in parts]
        total = total - total.mean(axis=0)
        return total, parts

    def density(self, psis, occ):
        """ElectronDensity.cpp:55-140."""
        nd, nk, N = self.grid.nd, len(self.kpts), self.nband
        spin_fac = 2.0 if self.nspin == 1 else 1.0
        rs = []
        for s in range(self.nspin):
            r = np.zeros(nd)
            for k in range(nk):
                f = occ[(s * nk + k) * N:(s * nk + k + 1) * N]
                w = np.where(f < 1e-16, 0.0, spin_fac * self.kw[k] * f)
                r += (np.abs(psis[s * nk + k]) ** 2) @ w
            rs.append(r)
        return [rs[0]] if self.nspin == 1 else [rs[0] + rs[1], rs[0], rs[1]]

    def initial_psis(self):
        """Wavefunction.cpp:53-76 with seed spin * 1000 + 1 (SCF.cpp:863-881)."""
        from . import rand_uniform
        nd, N = self.grid.nd, self.nband
        out = []
        for s in range(self.nspin):
            for k in self.kpts:
                if k is None:
                    v = rand_uniform(s * 1000 + 1, -0.5, 0.5, nd * N)
                    out.append(np.asfortranarray(v.reshape(N, nd).T))
                else:
                    v = rand_uniform(s * 1000 + 1, -0.5, 0.5, 2 * nd * N)
                    c = v[0::2] + 1j * v[1::2]
                    out.append(np.asfortranarray(c.reshape(N, nd).T))
        return out

    @staticmethod
    def initial_density(nspin, rho0, mag0=None):
        """SCF::set_initial_density (SCF.cpp:83-107): [total] for one spin,
        [total, up, down] for two with up/down = (total +- mag) / 2 floored
        at zero, or total / 2 each without mag."""
        t = np.array(rho0, float)
        if nspin == 1:
            return [t]
        if mag0 is None:
            return [t, 0.5 * t, 0.5 * t]
        m = np.asarray(mag0, float)
        return [t, np.maximum(0.5 * (t + m), 0.0), np.maximum(0.5 * (t - m), 0.0)]

    def run(self, callback=None, rho0=None, mag0=None):
        """The SCF loop (SCF.cpp:110-291) from the atomic densities, or from
        the restart density rho0 (total) and mag0 (two spins)."""
        nd = self.grid.nd
        dV = self.grid.dV
        E = self.E
        if rho0 is not None:
            rho = self.initial_density(self.nspin, rho0, mag0)
        else:
            rho = [E.rho_at.copy()] if self.nspin == 1 else [E.rho_at.copy(), 0.5 * E.rho_at, 0.5 * E.rho_at]
        veff = self.potential(rho)
        psis = self.initial_psis()
        err = 1.0
        spin_fac = 2.0 if self.nspin == 1 else 1.0
        self.history = []
        for it in range(self.max_iter):
            psis, eig = self.stage.step(it, veff, psis)
            occ = self.stage.occ
            new = self.density(psis, occ)
            # Energy.cpp:140-200 with the input density's fields
            w = self.stage.weights
            eband = float((w * occ * eig).sum())
            if self.smearing == 0:
                f = occ
                m = (f > 1e-14) & (f < 1 - 1e-14)
                S = float((w[m] * -(f[m] * np.log(f[m]) + (1 - f[m]) * np.log(1 - f[m]))).sum())
            else:
                x = self.beta * (eig - self.stage.fermi)
                S = float((w * (0.5 / math.sqrt(math.pi)) * np.exp(-x * x)).sum())
            entropy = -S / self.beta
            rt = rho[0]
            core = E.rho_core if E.has_nlcc else 0.0
            exc = float(((rt + core) * self.exc).sum() * dV)
            if self.nspin == 1:
                e2 = float((rt * self.vxc[0]).sum() * dV)
            else:
                e2 = float((rho[1] * self.vxc[0] + rho[2] * self.vxc[1]).sum() * dV)
            e3 = float((rt * self.phi).sum() * dV)
            ehart = 0.5 * float(((rt + E.b) * self.phi).sum() * dV)
            etot = eband - e2 + exc - e3 + ehart + E.eself + E.ec + entropy
            err = float(np.linalg.norm(new[0] - rt) / np.linalg.norm(new[0]))
            rec = dict(etotal=etot, eband=eband, exc=exc, ehart=ehart, entropy=entropy, error=err,
                       fermi=self.stage.fermi, eself=E.eself, ec=E.ec)
            self.history.append(rec)
            if callback:
                callback(it, rec)
            self.psis, self.occ, self.rho_in = psis, occ, rho
            if it >= self.min_iter and err < self.tol:
                rec["converged"] = True
                return rec
            if self.nspin == 1:
                rho = [self.mix(rho[0], new[0], 1)]
            else:
                x = self.mix(np.concatenate([rho[0], rho[1] - rho[2]]), np.concatenate([new[0], new[1] - new[2]]), 2)
                t, mg = x[:nd], x[nd:]
                rho = [t, 0.5 * (t + mg), 0.5 * (t - mg)]
            veff = self.potential(rho)
        rec = dict(self.history[-1])
        rec["converged"] = False
        return rec
