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
    preconditioner (KerkerPreconditioner.cpp:37-95, LinearSolver.cpp:23-130);
  * Pulay mixing with the density constraint (Mixer.cpp:116-266);
  * density, energies and the SCF loop (ElectronDensity.cpp:55-140,
    Energy.cpp:140-200, SCF.cpp:110-291, 470-790).
"""
from __future__ import annotations

import math

import numpy as np

from . import Grid, Lattice, Operator, positions_noncart, stencil_coeffs  # noqa: F401


# ---------------------------------------------------------------------------
# local pseudopotential set-up

RC_REF = 0.5   # cutoff of the smooth reference potential (Electrostatics.cpp:176)


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
