"""Pins of the oracle's forces (oracle/scf.py ScfOracle.forces,
Electrostatics.local_forces; Forces.cpp:82-420, 509-800, 1108-1288): the
force on an atom is minus the derivative of the converged SCF energy with
respect to its position.  Central differences of the total energy (SCF
converged to 1e-10, displacement +-0.002 Bohr) are compared with the
unsymmetrized sum of the local, nonlocal and core-correction terms, on a
grid fine enough (h about 0.25 Bohr for the hard C.pot) that the egg-box
error of the difference is below 1.5e-4 Ha/Bohr (3e-4 on the FCC cell at
h = 0.25 Bohr, 0.2 % of the force):

- orthogonal cell, Gamma point, PBE (real orbitals);
- two k-points, LDA (complex orbitals with Bloch phases);
- the FCC primitive cell (non-orthogonal: lattice-direction gradients,
  mixed-term pseudocharges and the map to Cartesian components);
- a model core charge (a Gaussian core density added to C.pot) for the
  core-correction term.
A wrong sign, a dropped factor 2 or 1/2, a missing term or a transposed
frame changes these forces by far more than the tolerance.  The occupation
factor of a spin-polarized run is pinned by a non-magnetic state: with two
spins converged to zero magnetization the forces equal the one-spin forces.
"""
import copy
import math

import numpy as np
import pytest

import oracle
import oracle.scf as osc
from systems import psp

KB = 3.1668115634556e-6
PAIR = np.array([[0.1, 0.12, 0.08], [0.38, 0.45, 0.33]])


def _scf(grid, ps, frac, **kw):
    O = osc.ScfOracle(grid, ps, [0, 0], frac, 8, 1 / (KB * 1000), tol=1e-10, solver="spectral", max_iter=300,
                      **kw)
    r = O.run()
    assert r["converged"]
    return O, r["etotal"]


def _check(grid, ps, frac, cart_dir, atom=1, d=0.002, tol=1.5e-4, **kw):
    O, _ = _scf(grid, ps, frac, **kw)
    F, parts = O.forces()
    raw = sum(parts)
    lat = grid.lattice
    # a Cartesian displacement d along cart_dir in fractional coordinates
    step = np.zeros(3)
    step[cart_dir] = d
    dfrac = step @ np.linalg.inv(lat.latvec)
    e = []
    for sgn in (1, -1):
        fr = frac.copy()
        fr[atom] += sgn * dfrac
        e.append(_scf(grid, ps, fr, **kw)[1])
    fd = -(e[0] - e[1]) / (2 * d)
    assert abs(raw[atom, cart_dir] - fd) < tol, (raw[atom, cart_dir], fd)
    # the terms are each large: the check is not carried by one of them
    assert abs(parts[0][atom, cart_dir]) > 0.01 and abs(parts[1][atom, cart_dir]) > 0.01
    assert np.abs(F.sum(axis=0)).max() < 1e-12
    return F, parts, fd


def _cube(n, a=6.0):
    return oracle.Grid(n, n, n, oracle.Lattice(np.diag([a] * 3), True))


def test_forces_orthogonal_gamma_pbe():
    _check(_cube(24), [oracle.load_psp8(psp("C.pot"))], PAIR, 0, xc=2)


def test_forces_kpoints():
    ks = [(0.0, 0.0, 0.0), (0.0, 0.0, math.pi / 6.0)]
    _check(_cube(24), [oracle.load_psp8(psp("C.pot"))], PAIR, 2, xc=0, kpts=ks, kpt_weights=[0.5, 0.5])


def test_forces_two_spins_equal_one_spin_without_magnetization():
    # Gaussian smearing at 20000 K keeps the pair non-magnetic
    g = _cube(16)
    ks = [(0.0, 0.0, 0.0), (0.0, 0.0, math.pi / 6.0)]
    kw = dict(xc=0, kpts=ks, kpt_weights=[0.5, 0.5], smearing=1)
    ps = [oracle.load_psp8(psp("C.pot"))]
    out = []
    for nspin in (1, 2):
        O = osc.ScfOracle(g, ps, [0, 0], PAIR, 8, 1 / (KB * 20000), tol=1e-9, solver="spectral", max_iter=300,
                          nspin=nspin, **kw)
        assert O.run()["converged"]
        if nspin == 2:
            rho = O.density(O.psis, O.occ)
            assert np.abs(rho[1] - rho[2]).max() < 1e-6
        out.append(O.forces()[0])
    assert np.abs(out[0]).max() > 0.05
    assert np.abs(out[0] - out[1]).max() < 1e-6


def test_forces_fcc_nonorthogonal():
    from systems import FCC_PRIMITIVE
    lv = 7.0 * FCC_PRIMITIVE
    g = oracle.Grid(20, 20, 20, oracle.Lattice(lv, False))
    frac = np.array([[0.02, 0.01, 0.03], [0.27, 0.24, 0.22]])
    _check(g, [oracle.load_psp8(psp("C.pot"))], frac, 0, xc=0, tol=3e-4)


def test_forces_core_correction():
    ps = copy.deepcopy(oracle.load_psp8(psp("C.pot")))
    ps.fchrg = 1.0
    ps.rho_c = 3.0 * np.exp(-(ps.r / 1.1) ** 2) / (4 * math.pi)
    ps.rho_c_d = oracle.spline_deriv(ps.r, ps.rho_c)
    F, parts, _ = _check(_cube(24), [ps], PAIR, 0, xc=0)
    assert abs(parts[2][1, 0]) > 5e-3
