"""Pins the CPU oracle before it is trusted as the checker.

Each test checks an oracle function against something independent of its own
formula: golden values held by the reference's tests, closed forms, brute
force on tiny inputs, or a textbook library routine (numpy/scipy LAPACK).
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from systems import DIAMOND, A_SI, psp

# ----------------------------------------------------------------- FD weights


def test_fd_weights_order6_exact():
    # tests/unit/test_FDStencil.cpp:7-57
    w1, w2 = oracle.fd_weights(6)
    assert np.allclose(w1, [0.0, 0.75, -0.15, 1.0 / 60.0], rtol=0, atol=1e-12)
    assert np.allclose(w2, [-(2.0 + 0.5 + 2.0 / 9.0), 1.5, -0.15, 2.0 * 0.05 / 9.0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("order", [2, 4, 6, 8, 12, 16])
def test_fd_weights_are_exact_on_polynomials(order):
    # a central scheme of order 2n differentiates x^k exactly for k <= 2n
    w1, w2 = oracle.fd_weights(order)
    n = order // 2
    p = np.arange(1, n + 1, dtype=float)
    for k in range(0, 2 * n + 1):
        t2 = w2[1:] * (p ** k + (-p) ** k)
        d2 = (w2[0] if k == 0 else 0.0) + np.sum(t2)
        assert abs(d2 - (2.0 if k == 2 else 0.0)) < 1e-14 * max(1.0, np.abs(t2).sum()), (order, k)
        t1 = w1[1:] * (p ** k - (-p) ** k)
        assert abs(np.sum(t1) - (1.0 if k == 1 else 0.0)) < 1e-14 * max(1.0, np.abs(t1).sum()), (order, k)


def test_fd_sum_rule_and_scaling():
    # test_FDStencil.cpp:59-105
    lat = oracle.Lattice(np.diag([10.0, 10.0, 10.0]), True)
    g = oracle.Grid(20, 20, 20, lat)
    st = oracle.stencil_coeffs(12, g)
    _, w2 = oracle.fd_weights(12)
    assert abs(w2[0] + 2 * w2[1:].sum()) < 1e-12
    assert np.allclose(st.D2[0], w2 / 0.25, rtol=0, atol=1e-10)
    assert st.max_eig > 0


def test_fd12_coefficients_match_sparc_dump():
    # Golden vector: DUMP_D2_COEFF of SPARC for BaTiO3_quick, 15^3 grid,
    # dx = 0.5086666... Bohr (tests/unit/test_CompareRef.cpp:165-200)
    ref = [-1.152800990777592e+01, 6.625469334751346e+00, -1.035229583554898e+00, 2.044897942824489e-01,
           -3.450765278516326e-02, 4.015435960455361e-03, -2.323747662300556e-04]
    L = 15 * 5.086666666666667e-01
    lat = oracle.Lattice(np.diag([L, L, L]), True)
    g = oracle.Grid(15, 15, 15, lat)
    st = oracle.stencil_coeffs(12, g)
    assert np.allclose(st.D2[0], ref, rtol=0, atol=1e-10)
    assert abs(g.h[0] - 5.086666666666667e-01) < 1e-14
    assert abs(g.dV - 1.316133176296297e-01) < 1e-12


def test_max_eigval_is_the_nyquist_eigenvalue():
    # (-1)^(i+j+k) is an eigenvector of the periodic FD Laplacian with the
    # largest eigenvalue; the oracle apply must return max_eig times it
    lat = oracle.Lattice(np.diag([7.0, 8.0, 9.0]), True)
    g = oracle.Grid(12, 14, 16, lat)
    op = oracle.Operator(g, 12)
    k, j, i = np.meshgrid(np.arange(16), np.arange(14), np.arange(12), indexing="ij")
    v = ((-1.0) ** (i + j + k)).ravel()
    y = op.apply(v[:, None], np.zeros(g.nd), nonlocal_=False)[:, 0]
    assert np.abs(y - op.st.max_eig * v).max() < 1e-10 * op.st.max_eig


# ------------------------------------------------------------------ halo wrap


def test_halo_extended_dimensions_and_wrap():
    # test_HaloExchange.cpp:12-75
    lat = oracle.Lattice(np.diag([10.0, 10.0, 10.0]), True)
    g = oracle.Grid(8, 9, 10, lat)
    op = oracle.Operator(g, 6)
    R = 3
    x = np.arange(g.nd, dtype=float)
    xe = op.halo(x).reshape(10 + 2 * R, 9 + 2 * R, 8 + 2 * R)
    ref = np.pad(x.reshape(10, 9, 8), R, mode="wrap")
    assert np.array_equal(xe, ref)


def test_halo_bloch_phases():
    # test_HaloExchange.cpp:78-176: k = 0.25 * 2pi/L gives e^{-i pi/2} = -i on
    # the low x ghost and +i on the high one; corners carry phase products
    L = 10.0
    lat = oracle.Lattice(np.diag([L, L, L]), True)
    n, R = 10, 3
    g = oracle.Grid(n, n, n, lat)
    kz = 0.125 * 2 * math.pi / L
    op = oracle.Operator(g, 6, kpt=(0.25 * 2 * math.pi / L, -0.3333 * 2 * math.pi / L, kz))
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    x = (i + 1.0 + 1j * (j + 0.1 * k)).ravel()
    xe = op.halo(x, complex_=True).reshape(n + 2 * R, n + 2 * R, n + 2 * R)
    X = x.reshape(n, n, n)
    assert xe[R + 1, R + 1, R - 1] == pytest.approx(X[1, 1, n - 1] * (-1j), abs=1e-12)
    assert xe[R + 1, R + 1, R + n] == pytest.approx(X[1, 1, 0] * (1j), abs=1e-12)
    pz = np.exp(-1j * kz * L)
    assert xe[R - 1, R + 2, R + 3] == pytest.approx(X[n - 1, 2, 3] * pz, abs=1e-12)
    # Gamma: complex halo equals the real one (test_HaloExchange.cpp:178-219)
    op0 = oracle.Operator(g, 12)
    xr = np.sin(0.1 * np.arange(g.nd))
    assert np.allclose(op0.halo(xr.astype(complex), complex_=True).real, op0.halo(xr), atol=1e-14)


# ------------------------------------------------------- local Hamiltonian


def test_free_particle_sin():
    # test_Hamiltonian.cpp:39-72: -1/2 lap sin(x) = 1/2 sin(x), FD12, 40^3
    L = 2 * math.pi
    g = oracle.Grid(40, 40, 40, oracle.Lattice(np.diag([L, L, L]), True))
    op = oracle.Operator(g, 12)
    x = np.arange(40) * g.h[0]
    psi = np.tile(np.sin(x), 40 * 40)
    y = op.apply(psi[:, None], np.zeros(g.nd), nonlocal_=False)[:, 0]
    assert np.abs(y - 0.5 * psi).max() < 1e-6


def test_constant_gives_potential():
    # test_Hamiltonian.cpp:74-112
    g = oracle.Grid(20, 20, 20, oracle.Lattice(np.diag([2 * math.pi] * 3), True))
    op = oracle.Operator(g, 12)
    V = 0.5 * (np.arange(g.nd) % 10)
    y = op.apply(np.full((g.nd, 2), 2.0), V, nonlocal_=False)
    assert np.abs(y - 2.0 * V[:, None]).max() < 1e-10


def _symbol(st, q, h):
    s = 0.0
    for a in range(3):
        s += st.D2[a, 0] + 2 * sum(st.D2[a, p] * math.cos(q[a] * p * h[a]) for p in range(1, st.fdn + 1))
    return -0.5 * s


def test_plane_wave_symbol_with_bloch_phase():
    # e^{i(k+G).r} is an exact eigenvector of the Bloch FD operator with
    # eigenvalue given by the stencil symbol at k+G (closed form)
    L = np.array([9.0, 10.0, 11.0])
    lat = oracle.Lattice(np.diag(L), True)
    n = (12, 10, 14)
    g = oracle.Grid(*n, lat)
    kpt = 2 * math.pi * np.array([0.25, -0.125, 0.375]) / L
    op = oracle.Operator(g, 12, kpt=tuple(kpt))
    G = 2 * math.pi * np.array([1, 2, -1]) / L
    q = kpt + G
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    r = np.stack([i * g.h[0], j * g.h[1], k * g.h[2]], axis=-1).reshape(-1, 3)
    psi = np.exp(1j * r @ q)
    y = op.apply(psi[:, None], np.zeros(g.nd), nonlocal_=False)[:, 0]
    lam = _symbol(op.st, q, g.h)
    assert np.abs(y - lam * psi).max() < 1e-10 * abs(lam)


FCC = 0.5 * A_SI * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])


def _nonorth_symbol(st, phi):
    # FD symbol of the non-orthogonal operator on exp(i sum_a phi_a n_a):
    # second-difference terms give cosines; each mixed term
    # sum_pq c[p] d1[q] (x++ - x+- - x-+ + x--) gives -4 sum_p c[p] sin(p phi_a)
    # * sum_q d1[q] sin(q phi_b)
    R = st.fdn
    s = 0.0
    for a in range(3):
        s += st.D2[a, 0] + 2 * sum(st.D2[a, p] * math.cos(p * phi[a]) for p in range(1, R + 1))
    sn = lambda c, f: sum(c[p] * math.sin(p * f) for p in range(1, R + 1))  # noqa: E731
    s -= 4 * sn(st.D2m[0], phi[0]) * sn(st.D1[1], phi[1])
    s -= 4 * sn(st.D2m[1], phi[0]) * sn(st.D1[2], phi[2])
    s -= 4 * sn(st.D2m[2], phi[1]) * sn(st.D1[2], phi[2])
    return -0.5 * s


def test_nonorth_plane_wave_symbol_with_bloch_phase():
    # FCC primitive cell, unequal grid dimensions so a swapped axis shows:
    # a Bloch plane wave is an eigenvector with the stencil symbol as its
    # eigenvalue (closed form, independent of the real-space loops)
    lat = oracle.Lattice(FCC, False)
    n = (12, 14, 16)
    g = oracle.Grid(*n, lat)
    L = lat.lengths
    kpt = np.array([0.25, -0.125, 0.375]) * 2 * math.pi / L   # theta_a = kpt_a L_a
    op = oracle.Operator(g, 12, kpt=tuple(kpt))
    m = np.array([1, -2, 3])
    phi = (kpt * L + 2 * math.pi * m) / np.array(n)
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    idx = np.stack([i, j, k], axis=-1).reshape(-1, 3)
    psi = np.exp(1j * idx @ phi)
    y = op.apply(psi[:, None], np.zeros(g.nd), nonlocal_=False, complex_=True)[:, 0]
    lam = _nonorth_symbol(op.st, phi)
    assert np.abs(y - lam * psi).max() < 1e-10 * abs(lam)
    assert abs(op.st.D2m[0, 1]) > 0 and abs(op.st.D2m[2, 1]) > 0


@pytest.mark.parametrize("m", [(1, 0, 0), (0, 1, -1), (1, 1, 1), (2, -1, 1)])
def test_nonorth_laplacian_continuum_limit(m):
    # -1/2 lap cos(G.r) = 1/2 |G|^2 cos(G.r) for G = 2 pi sum_a m_a b_a; at
    # long wavelength FD12 is exact to ~1e-7, so the lapc_T weights, the
    # factor 2 of the mixed terms and their 1/h scaling are all pinned
    lat = oracle.Lattice(FCC, False)
    n = (30, 32, 34)
    g = oracle.Grid(*n, lat)
    op = oracle.Operator(g, 12)
    B = np.linalg.inv(FCC).T                      # rows b_a with a_i . b_j = delta_ij
    G = 2 * math.pi * np.asarray(m, float) @ B
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    frac = np.stack([i / n[0], j / n[1], k / n[2]], axis=-1).reshape(-1, 3)
    r = frac @ FCC
    psi = np.cos(r @ G)
    y = op.apply(psi[:, None], np.zeros(g.nd), nonlocal_=False)[:, 0]
    want = 0.5 * (G @ G) * psi
    assert np.abs(y - want).max() < 2e-6 * 0.5 * (G @ G)


def test_nonorth_sphere_membership_brute_force():
    # every (atom, image, point) with Cartesian distance <= rc, found by brute
    # force over +-1 images in an FCC primitive cell
    lat = oracle.Lattice(FCC, False)
    n = 12
    g = oracle.Grid(n, n, n, lat)
    ps = [oracle.load_psp8(psp("C.pot"))]
    frac = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    op = oracle.setup_system(g, ps, [0, 0], frac, 12)
    rc = ps[0].rc_max
    L = lat.lengths
    got = {(im.atom, tuple(np.round(im.shift / L).astype(int)), int(p)) for im in op.images for p in im.gpos}
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    fr = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) / n
    want = set()
    rng = range(-2, 3)
    for a, f in enumerate(frac):
        for sx in rng:
            for sy in rng:
                for sz in rng:
                    sh = np.array([sx, sy, sz])
                    d = np.linalg.norm((fr - (f + sh)) @ FCC, axis=1)
                    for p in np.nonzero(d <= rc)[0]:
                        want.add((a, (sx, sy, sz), int(p)))
    assert len(want) > 0 and got == want


# ---------------------------------------------------------------- psp8 and splines


def test_psp8_reads_file_values():
    p = oracle.load_psp8(psp("C.pot"))
    assert (p.zion, p.lmax, p.lloc, p.mmax, p.ppl) == (4.0, 1, 4, 600, [2, 2])
    assert p.nproj_per_atom == 8
    assert p.gamma[(0, 0)] == pytest.approx(8.1302825364155)
    assert p.gamma[(1, 1)] == pytest.approx(-0.85555610690489)
    rows = []
    with open(psp("C.pot")) as f:
        lines = f.read().splitlines()
    for ln in lines[7:7 + 600]:
        rows.append([float(t.replace("D", "E")) for t in ln.split()])
    rows = np.array(rows)
    r = rows[:, 1]
    assert np.array_equal(p.r, r)
    u = rows[1:, 2] / r[1:]
    assert np.allclose(p.udv[(0, 0)][1:], u, rtol=1e-15, atol=0)
    assert p.udv[(0, 0)][0] == p.udv[(0, 0)][1]          # r = 0 copy rule
    nz = np.nonzero(np.abs(p.udv[(0, 0)]) > 1e-12)[0]
    assert p.rc[0] == r[nz[-1]]


def test_spline_interpolates_nodes_and_smooth_functions():
    x = np.cumsum(np.r_[0.0, 0.05 + 0.02 * np.sin(np.arange(59))])
    # the end conditions reproduce cubics exactly: derivatives and values
    yc = 1 + 0.5 * x - 0.3 * x ** 2 + 0.2 * x ** 3
    ydc = oracle.spline_deriv(x, yc)
    assert np.abs(ydc - (0.5 - 0.6 * x + 0.6 * x ** 2)).max() < 1e-12
    xm = 0.5 * (x[:-1] + x[1:])
    fc = oracle.spline_eval(x, yc, ydc, xm)
    assert np.abs(fc - (1 + 0.5 * xm - 0.3 * xm ** 2 + 0.2 * xm ** 3)).max() < 1e-13
    # uniform table (the psp8 case): direct interval lookup, same exactness
    xu = np.arange(40) * 0.01
    yu = 2 - xu + 3 * xu ** 3
    fu = oracle.spline_eval(xu, yu, oracle.spline_deriv(xu, yu), xu[:-1] + 0.0037)
    assert np.abs(fu - (2 - (xu[:-1] + 0.0037) + 3 * (xu[:-1] + 0.0037) ** 3)).max() < 1e-13
    y = np.exp(-x) * np.cos(3 * x)
    yd = oracle.spline_deriv(x, y)
    assert np.allclose(oracle.spline_eval(x, y, yd, x[:-1]), y[:-1], atol=1e-15)
    # clamping outside the table
    assert oracle.spline_eval(x, y, yd, np.array([-1.0, x[-1] + 1]))[0] == y[0]


# ------------------------------------------------------------- spherical harmonics


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = math.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], axis=1)


@pytest.mark.parametrize("l", [0, 1, 2, 3])
def test_ylm_addition_theorem(l):
    # sum_m Y_lm(u) Y_lm(v) = (2l+1)/(4 pi) P_l(u.v) fixes normalisation and
    # completeness of the real basis independently of the m convention
    u = _fib_sphere(40)
    v = _fib_sphere(37)[::-1] * 2.5
    ru = np.ones(len(u)); rv = np.full(len(v), 2.5)
    su = np.stack([oracle.ylm_real(l, m, *u.T, ru) for m in range(-l, l + 1)])
    sv = np.stack([oracle.ylm_real(l, m, *v.T, rv) for m in range(-l, l + 1)])
    lhs = su.T @ sv
    cosg = (u @ (v / 2.5).T)
    Pl = np.polynomial.legendre.legval(cosg, [0] * l + [1])
    assert np.abs(lhs - (2 * l + 1) / (4 * math.pi) * Pl).max() < 1e-12


# --------------------------------------------------- influence spheres and Chi


def _si8(n):
    lat = oracle.Lattice(np.diag([A_SI] * 3), True)
    g = oracle.Grid(n, n, n, lat)
    ps = [oracle.load_psp8(psp("C.pot"))]
    op = oracle.setup_system(g, ps, [0] * 8, DIAMOND, 12)
    return g, ps, op


def test_sphere_membership_brute_force():
    g, ps, op = _si8(16)
    rc = ps[0].rc_max
    got = {(im.atom, tuple(im.shift), int(p)) for im in op.images for p in im.gpos}
    want = set()
    k, j, i = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    r = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1) * g.h
    for a, f in enumerate(DIAMOND * A_SI):
        for sx in (-1, 0, 1):
            for sy in (-1, 0, 1):
                for sz in (-1, 0, 1):
                    sh = np.array([sx, sy, sz]) * A_SI
                    d = np.linalg.norm(r - (f + sh), axis=1)
                    for p in np.nonzero(d <= rc)[0]:
                        want.add((a, tuple(sh), int(p)))
    assert got == want


def test_chi_obeys_addition_theorem():
    g, ps, op = _si8(16)
    p = ps[0]
    for im in op.images[:4]:
        gg = im.gpos
        i = gg % 16; j = (gg // 16) % 16; k = gg // 256
        c = DIAMOND[im.atom] * A_SI + im.shift
        d = np.stack([i * g.h[0] - c[0], j * g.h[1] - c[1], k * g.h[2] - c[2]], axis=1)
        r = np.linalg.norm(d, axis=1)
        cols = 0
        for l in range(p.lmax + 1):
            for pp in range(p.ppl[l]):
                blk = im.chi[:, cols:cols + 2 * l + 1]
                cols += 2 * l + 1
                u = oracle.spline_eval(p.r, p.udv[(l, pp)], p.udv_d[(l, pp)], r)
                rr = np.where(r > 0, r, 1.0)
                cosg = np.clip((d / rr[:, None]) @ (d / rr[:, None]).T, -1, 1)
                Pl = np.polynomial.legendre.legval(cosg, [0] * l + [1])
                want = (2 * l + 1) / (4 * math.pi) * np.outer(u, u) * Pl
                if l > 0:
                    want[r == 0, :] = 0; want[:, r == 0] = 0
                assert np.abs(blk @ blk.T - want).max() < 1e-11 * max(1, np.abs(want).max())


def test_nonlocal_apply_matches_dense_operator():
    g, ps, op = _si8(16)
    nd = g.nd
    ip, gam = oracle.atom_gamma(ps, [0] * 8)
    M = np.zeros((8, nd, 8))
    for im in op.images:
        np.add.at(M[im.atom], im.gpos, im.chi)
    Vnl = sum(g.dV * M[a] @ np.diag(gam[ip[a]:ip[a + 1]]) @ M[a].T for a in range(8))
    assert np.abs(Vnl - Vnl.T).max() < 1e-12 * np.abs(Vnl).max()
    X = np.random.default_rng(0).uniform(-0.5, 0.5, (nd, 3))
    full = op.apply(X, np.zeros(nd))
    loc = op.apply(X, np.zeros(nd), nonlocal_=False)
    assert np.abs((full - loc) - Vnl @ X).max() < 1e-11 * np.abs(Vnl @ X).max()
    # k-point: V_nl(k) = dV sum_a M_a(k) G M_a(k)^H with M_a(k) = sum_img e^{+ik.R} P^T Chi
    kpt = (0.3, -0.2, 0.45)
    op.set_kpoint(kpt)
    Mk = np.zeros((8, nd, 8), complex)
    for im in op.images:
        np.add.at(Mk[im.atom], im.gpos, np.exp(1j * np.dot(kpt, im.shift)) * im.chi)
    Vk = sum(g.dV * Mk[a] @ np.diag(gam[ip[a]:ip[a + 1]]) @ Mk[a].conj().T for a in range(8))
    Xc = X + 1j * X[:, ::-1]
    fullk = op.apply(Xc, np.zeros(nd), complex_=True)
    lock = op.apply(Xc, np.zeros(nd), nonlocal_=False, complex_=True)
    assert np.abs((fullk - lock) - Vk @ Xc).max() < 1e-11 * np.abs(Vk @ Xc).max()


# ---------------------------------------------------- filter, Lanczos, solve


def _dense(op, nd, veff):
    return op.apply(np.eye(nd), veff)


def test_chebyshev_filter_closed_form():
    # on eigenvectors Y = p_m(lambda) X, p_m(l) = T_m((l-c)/e) / T_m((lmin-c)/e)
    g = oracle.Grid(10, 10, 10, oracle.Lattice(np.diag([6.0] * 3), True))
    op = oracle.Operator(g, 12)
    veff = np.random.default_rng(1).uniform(-1, 0, g.nd)
    Hd = _dense(op, g.nd, veff)
    w, Q = np.linalg.eigh(0.5 * (Hd + Hd.T))
    X = Q[:, :6]
    lmin, lc, lmax = w[0] - 0.1, w[20], w[-1] * 1.01
    e = (lmax - lc) / 2; c = (lmax + lc) / 2
    for m in (1, 2, 5, 20):
        Y = op.cheb_filter(X, veff, lc, lmin, lmax, m)
        Tm = np.polynomial.chebyshev.chebval((w[:6] - c) / e, [0] * m + [1])
        T0 = np.polynomial.chebyshev.chebval((lmin - c) / e, [0] * m + [1])
        assert np.abs(Y - X * (Tm / T0)).max() < 1e-10, m


def test_tridiagonal_extremes_match_lapack():
    import ctypes
    rng = np.random.default_rng(3)
    for n in (2, 5, 40):
        d = rng.normal(size=n); e = rng.normal(size=n)
        mn = ctypes.c_double(); mx = ctypes.c_double()
        oracle.lib().orc_tridiag_extremes(oracle._p(d), oracle._p(e), n, ctypes.byref(mn), ctypes.byref(mx))
        w = sla.eigvalsh_tridiagonal(d, e[:n - 1])
        assert abs(mn.value - w[0]) < 1e-13 * max(1, abs(w[0]))
        assert abs(mx.value - w[-1]) < 1e-13 * max(1, abs(w[-1]))


def test_glibc_rand_stream():
    # glibc TYPE_3 generator after srand(1): first output 1804289383
    u = oracle.rand_uniform(1, 0.0, 1.0, 2)
    assert u[0] == 1804289383 / 2147483647
    assert u[1] == 846930886 / 2147483647


def test_lanczos_brackets_the_spectrum():
    g, ps, op = _si8(12)
    veff = np.random.default_rng(2).uniform(-1, 0, g.nd)
    Hd = _dense(op, g.nd, veff)
    w = np.linalg.eigvalsh(0.5 * (Hd + Hd.T))
    mn, mx, it = op.lanczos_bounds(veff, tol=1e-2, max_iter=1000)
    assert w[0] - 1e-9 <= mn + 0.1 <= w[0] + 0.05 * abs(w[0]) + 1e-2
    assert w[-1] * (1 - 0.02) <= mx / 1.01 <= w[-1] + 1e-9
    assert 0 < it < 1000


def test_solve_converges_to_lowest_eigenvalues():
    g, ps, op = _si8(12)
    veff = np.random.default_rng(4).uniform(-1, 0, g.nd)
    Hd = _dense(op, g.nd, veff)
    w = np.linalg.eigvalsh(0.5 * (Hd + Hd.T))
    mn, mx, _ = op.lanczos_bounds(veff)
    lc = 0.5 * (mn + mx)
    X = np.asfortranarray(np.random.default_rng(5).uniform(-0.5, 0.5, (g.nd, 10)))
    for _ in range(25):
        X, eig, lc = op.solve(X, veff, lc, mn, mx, 20)
    assert np.abs(eig[:6] - w[:6]).max() < 1e-8


# ---------------------------------------------------------------- k-points


def test_kpoints_fcc_6x6x6_reduces_to_112():
    a = A_SI
    lat = oracle.Lattice(0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]]), False)
    cart, red, w = oracle.kpoints((6, 6, 6), (0, 0, 0), lat)
    assert len(cart) == 112
    assert abs(w.sum() - 1.0) < 1e-12
    assert np.sum(w == 1.0 / 216) == 8


# ------------------------------------------------------------- occupations


@pytest.mark.parametrize("kind", [0, 1])
def test_fermi_level_closed_forms(kind):
    beta = 315.77
    # f(x) + f(-x) = 1 for both smearings: levels symmetric about 0 with
    # half filling put Ef at 0
    e = np.array([-0.7, -0.2, 0.2, 0.7])
    assert abs(oracle.fermi_level(e, [2.0] * 4, 4.0, beta, kind)) < 1e-12
    # one level, half occupied: Ef at the level
    assert abs(oracle.fermi_level([0.3], [2.0], 1.0, beta, kind) - 0.3) < 1e-12
    # occupations sum to Ne
    e = np.linspace(-0.5, 0.4, 17)
    w = np.full(17, 2.0)
    ef = oracle.fermi_level(e, w, 19.0, 50.0, kind)
    assert abs((w * oracle.smearing(e - ef, 50.0, kind)).sum() - 19.0) < 1e-11


# ------------------------------------------------------------ eigen stage


def test_eigen_stage_bookkeeping_two_spins():
    # SCF::estimate_spectral_bounds + solve_eigenproblem (SCF.cpp:310-467,
    # 883-935) with two spins whose potentials differ by a constant c:
    # - H + c has the same Krylov space, so its Lanczos run takes the same
    #   steps with every eigenvalue estimate shifted by c (closed form);
    # - spin 1's last eigenvalue is then the larger one and must set lambda_c
    #   (a lambda_c taken from spin 0 alone fails);
    # - eig_min keeps the Lanczos value in the first SCF iteration and takes
    #   eig(s, k0)[0] afterwards; eig_max never changes;
    # - rho_trigger passes in iteration 0, nchefsi later, each over every spin;
    # - the occupations hold Ne electrons;
    # - once converged, spin 1's eigenvalues are spin 0's plus c.
    g, ps, op = _si8(12)
    c = 0.05
    v0 = np.random.default_rng(7).uniform(-1, 0, g.nd)
    veff = np.concatenate([v0, v0 + c])
    nb, ne, beta = 8, 10.0, 1.0 / (3.166811563e-6 * 1000.0)
    mn0, mx0, it0 = op.lanczos_bounds(v0)
    mn1, mx1, it1 = op.lanczos_bounds(v0 + c)
    assert it1 == it0
    assert abs((mn1 + 0.1) - (mn0 + 0.1) - c) < 1e-9
    assert abs(mx1 / 1.01 - mx0 / 1.01 - c) < 1e-9
    st = oracle.ScfEigStage(op, nb, ne, beta, nspin=2, rho_trigger=3, nchefsi=2)
    X = np.asfortranarray(np.random.default_rng(8).uniform(-0.5, 0.5, (g.nd, nb)))
    psis = [X.copy(order="F"), X.copy(order="F")]
    psis, e = st.step(0, veff, psis)
    assert st.solves == 3 * 2
    assert list(st.eig_min) == [mn0, mn1] and list(st.eig_max) == [mx0, mx1]
    e = e.reshape(2, nb)
    assert e[1, -1] > e[0, -1] + 0.01
    assert abs(st.lambda_c - (e[1, -1] + 0.1)) < 1e-15
    assert abs((st.weights * st.occ).sum() - ne) < 1e-10
    for it in range(1, 6):
        psis, e = st.step(it, veff, psis)
        e = e.reshape(2, nb)
        assert st.solves == 6 + 4 * it
        assert list(st.eig_min) == [e[0, 0], e[1, 0]]
        assert list(st.eig_max) == [mx0, mx1]
        assert abs(st.lambda_c - (max(e[0, -1], e[1, -1]) + 0.1)) < 1e-15
    assert np.abs(e[1, :5] - e[0, :5] - c).max() < 1e-6
