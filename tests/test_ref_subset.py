"""Oracle restatement versus the reference's own code, for the leaf steps that
build from the reference sources alone (oracle/_ref, compiled by
oracle/ref_build.py in the container that holds the reference).  Skipped
where that library is absent."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from systems import A_SI, ROOT, psp

REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (no reference tree)")


@pytest.fixture(scope="module")
def ref():
    return ctypes.CDLL(REF)


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("order", [6, 12])
@pytest.mark.parametrize("latvec,orth", [
    (np.diag([10.26, 9.3, 12.1]), 1),
    (0.5 * A_SI * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]]), 0),
    (np.array([[7.57, 0, 0], [0.4, 7.6, 0], [0.3, -0.2, 8.0]]), 0),
])
def test_stencil_matches_reference(ref, order, latvec, orth):
    m = order // 2 + 1
    lv = np.ascontiguousarray(latvec, float)
    w1, w2 = np.zeros(m), np.zeros(m)
    D1, D2, D2m = np.zeros(3 * m), np.zeros(3 * m), np.zeros(3 * m)
    me, dV = ctypes.c_double(), ctypes.c_double()
    h = np.zeros(3)
    assert ref.ref_stencil(order, 20, 24, 28, P(lv), orth, P(w1), P(w2), P(D1), P(D2), P(D2m), ctypes.byref(me),
                           P(h), ctypes.byref(dV)) == 0
    lat = oracle.Lattice(lv, bool(orth))
    g = oracle.Grid(20, 24, 28, lat)
    st = oracle.stencil_coeffs(order, g)
    rw1, rw2 = oracle.fd_weights(order)
    assert np.array_equal(w1, rw1) and np.array_equal(w2, rw2)
    assert np.allclose(D1.reshape(3, m), st.D1, rtol=1e-15, atol=0)
    assert np.allclose(D2.reshape(3, m), st.D2, rtol=1e-13, atol=0)
    assert np.allclose(D2m.reshape(3, m), st.D2m, rtol=1e-13, atol=1e-300)
    assert abs(me.value - st.max_eig) <= 1e-13 * st.max_eig
    assert np.allclose(h, g.h, rtol=1e-15) and abs(dV.value - g.dV) <= 1e-14 * g.dV


@pytest.mark.parametrize("cplx", [0, 1])
def test_halo_matches_reference(ref, cplx):
    n = (9, 8, 10)
    L = np.array([9.0, 8.5, 11.0])
    kpt = 2 * np.pi * np.array([0.25, -0.3333, 0.125]) / L
    lat = oracle.Lattice(np.diag(L), True)
    g = oracle.Grid(*n, lat)
    op = oracle.Operator(g, 12, kpt=tuple(kpt) if cplx else (0, 0, 0))
    rng = np.random.default_rng(0)
    x = rng.normal(size=g.nd * 2)
    if cplx:
        x = x + 1j * rng.normal(size=g.nd * 2)
    mine = op.halo(x, complex_=bool(cplx))
    out = np.zeros_like(mine)
    xin = np.ascontiguousarray(x)
    assert ref.ref_halo(n[0], n[1], n[2], 6, 2, P(xin), P(out), cplx, P(np.ascontiguousarray(kpt)), P(L)) == 0
    assert np.abs(out - mine).max() <= 1e-15 * np.abs(out).max()


def test_psp8_matches_reference(ref):
    sizes = np.zeros(5, np.int32)
    zion = ctypes.c_double()
    r = np.zeros(600)
    tabs = np.zeros(2 * 8 * 600)
    gam = np.zeros(8)
    rc = np.zeros(8)
    assert ref.ref_psp8(psp("C.pot").encode(), P(sizes), ctypes.byref(zion), P(r), P(tabs), P(gam), P(rc), 8) == 0
    lmax, lloc, mmax, nproj, nt = sizes
    p = oracle.load_psp8(psp("C.pot"))
    assert (lmax, lloc, mmax, nproj) == (p.lmax, p.lloc, p.mmax, p.nproj_per_atom) and zion.value == p.zion
    assert np.array_equal(r[:mmax], p.r)
    t = 0
    for l in range(lmax + 1):
        if l == lloc:
            continue
        assert rc[l] == p.rc[l]
        for q in range(p.ppl[l]):
            assert np.array_equal(tabs[2 * t * mmax:(2 * t + 1) * mmax], p.udv[(l, q)])
            d = tabs[(2 * t + 1) * mmax:(2 * t + 2) * mmax]
            assert np.abs(d - p.udv_d[(l, q)]).max() <= 1e-12 * np.abs(d).max()
            assert gam[t] == p.gamma[(l, q)]
            t += 1
    assert t == nt


def test_spline_matches_reference(ref):
    x = np.cumsum(np.r_[0.0, 0.05 + 0.02 * np.sin(np.arange(79))])
    y = np.exp(-x) * np.cos(3 * x)
    xi = np.linspace(-0.1, x[-1] + 0.1, 333)
    yd = np.zeros_like(x)
    yi = np.zeros_like(xi)
    ref.ref_spline(P(x), P(y), len(x), P(xi), len(xi), P(yd), P(yi))
    myd = oracle.spline_deriv(x, y)
    assert np.abs(yd - myd).max() <= 1e-13 * np.abs(yd).max()
    assert np.abs(yi - oracle.spline_eval(x, y, myd, xi)).max() <= 1e-14


@pytest.mark.parametrize("K,shift,latvec,orth", [
    ((6, 6, 6), (0, 0, 0), 0.5 * A_SI * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]]), 0),
    ((2, 2, 2), (0.5, 0.5, 0.5), np.diag([A_SI] * 3), 1),
    ((3, 4, 5), (0, 0.5, 0), np.diag([7.0, 8.0, 9.0]), 1),
])
def test_kpoints_match_reference(ref, K, shift, latvec, orth):
    lv = np.ascontiguousarray(latvec, float)
    kc = np.zeros(3 * 512)
    w = np.zeros(512)
    n = ref.ref_kpoints(K[0], K[1], K[2], P(np.array(shift, float)), P(lv), orth, P(kc), P(w), 512)
    cart, red, nw = oracle.kpoints(K, shift, oracle.Lattice(lv, bool(orth)))
    assert n == len(cart)
    assert np.allclose(kc[:3 * n].reshape(n, 3), cart, rtol=0, atol=1e-14)
    assert np.allclose(w[:n] / np.prod(K), nw, rtol=0, atol=1e-15)


@pytest.mark.parametrize("cplx", [0, 1])
def test_lanczos_start_vector_matches_reference(ref, cplx):
    # the reference's random_init as g++ compiles it: for complex entries the
    # two rand() calls of Complex(rand(), rand()) run imaginary part first
    n = 257
    out = np.zeros(2 * n if cplx else n)
    ref.ref_lanczos_start(cplx, n, 1, P(out))
    want = out[0::2] + 1j * out[1::2] if cplx else out
    got = oracle.lanczos_start(cplx, n, 1)
    assert np.array_equal(got, want)


_MATVEC = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("cplx", [0, 1])
def test_lanczos_matches_reference(ref, cplx):
    # the reference's own lanczos_bounds<T> (Lanczos.hpp:104-198, LAPACK
    # dsterf_ from the MKL runtime) run with the oracle's H.v as its matvec:
    # the oracle's Lanczos (its own recurrence, start vector and tridiagonal
    # extremes) must stop after the same number of steps with the same bounds
    from systems import si8
    k = (0.25 * 2 * np.pi / A_SI, -0.125 * 2 * np.pi / A_SI, 0.375 * 2 * np.pi / A_SI) if cplx else (0, 0, 0)
    s = si8(14, kpt=k)
    op = s.oracle_op()
    nd = s.nd

    def mv(x, y):
        n = 2 * nd if cplx else nd
        xa = np.ctypeslib.as_array(x, (n,))
        ya = np.ctypeslib.as_array(y, (n,))
        v = xa.view(np.complex128) if cplx else xa
        out = op.apply(v.reshape(nd, 1), s.veff, complex_=bool(cplx)).ravel()
        ya[:] = out.view(np.float64) if cplx else out

    cb = _MATVEC(mv)
    for tol, max_iter in ((1e-2, 1000), (1e-6, 1000), (0.0, 7)):
        mn, mx, nmv = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        ref.ref_lanczos(cplx, nd, cb, ctypes.c_double(tol), max_iter, 1, ctypes.byref(mn), ctypes.byref(mx),
                        ctypes.byref(nmv))
        omn, omx, it = op.lanczos_bounds(s.veff, tol=tol, max_iter=max_iter)
        assert it == nmv.value - 1, (tol, it, nmv.value)
        assert abs(omn - mn.value) <= 1e-10 * max(1.0, abs(mn.value)), (tol, omn, mn.value)
        assert abs(omx - mx.value) <= 1e-10 * abs(mx.value), (tol, omx, mx.value)


@pytest.mark.parametrize("cplx", [0, 1])
def test_randomize_matches_reference(ref, cplx):
    # Wavefunction::randomize / randomize_kpt (Wavefunction.cpp:53-76) against
    # the library's lynx_randomize (through the pylynx-style Wavefunction)
    # and the oracle's glibc stream, bitwise
    from lynx_amd import core
    nd, nb, seed = 1013, 5, 1001
    out = np.zeros(2 * nd * nb if cplx else nd * nb)
    ref.ref_randomize(cplx, nd, nb, ctypes.c_uint(seed), P(out))
    want = (out[0::2] + 1j * out[1::2] if cplx else out).reshape(nb, nd).T
    w = core.Wavefunction().allocate(nd, nb, 2, 1, bool(cplx))
    w.randomize(1, 0, seed)
    assert np.array_equal(w.psi(1, 0), want)
    assert not np.any(w.psi(0, 0))
    stream = oracle.rand_uniform(seed, -0.5, 0.5, out.size)
    assert np.array_equal(stream, out)


@pytest.mark.parametrize("gapped", [False, True])
@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("nspin,nkpt", [(1, 1), (2, 3)])
def test_occupation_matches_reference(ref, kind, nspin, nkpt, gapped):
    # Occupation::compute (Occupation.cpp:86-176) against the oracle's Fermi
    # level and the library's host lynx_fermi_level.  In a gapped spectrum
    # (15 electrons filling the lowest levels exactly, a 0.3 Ha gap above)
    # the root is flat across the gap and the point returned is fixed only
    # by the root finder's own steps.
    from lynx_amd import _lib as L
    nb = 13
    rng = np.random.default_rng(11 + kind + nspin)
    eig = np.sort(rng.uniform(-0.4, 0.3, (nspin, nkpt, nb)), axis=2)
    if gapped:
        eig = np.sort(rng.uniform(-0.4, -0.1, (nspin, nkpt, nb)), axis=2)
        if nspin == 1:
            eig[:, :, 7:] += 0.3
        else:
            eig[0, :, 8:] += 0.3
            eig[1, :, 7:] += 0.3
    eig = eig.ravel()
    kw = rng.uniform(0.5, 1.5, nkpt)
    kw /= kw.sum()
    ne, beta = 15.0, 1.0 / (3.166811563e-6 * 3000.0)
    rocc = np.zeros_like(eig)
    ref.ref_occupation.restype = ctypes.c_double
    ref_ef = ref.ref_occupation(nspin, nkpt, nb, P(eig), P(kw), ctypes.c_double(ne), ctypes.c_double(beta), kind,
                                P(rocc))
    spin_fac = 2.0 if nspin == 1 else 1.0
    w = np.concatenate([np.full(nb, spin_fac * kw[k]) for s in range(nspin) for k in range(nkpt)])
    ef = oracle.fermi_level(eig, w, ne, beta, kind)
    assert abs(ef - ref_ef) < 1e-12
    assert np.abs(oracle.smearing(eig - ef, beta, kind) - rocc).max() < 1e-10
    occ = np.zeros_like(eig)
    lef = ctypes.c_double()
    L.check(L.lib().lynx_fermi_level(nspin, nkpt, nb, L.dptr(eig), L.dptr(kw), ne, beta, kind, L.dptr(occ),
                                     ctypes.byref(lef)))
    assert abs(lef.value - ref_ef) < 1e-12
    assert np.abs(occ - rocc).max() < 1e-10


@pytest.mark.parametrize("case", ["si8_orth", "si2_fcc", "becu_two_species"])
def test_electrostatics_matches_reference(ref, case):
    # Crystal::compute_atom_influence + Electrostatics (pseudocharge, b_ref,
    # Vc, Eself, Ec, atomic density; Electrostatics.cpp:139-463) compiled
    # from source, against the oracle and the library's host set-up
    import oracle.scf as osc
    from lynx_amd import _lib as L
    from systems import DIAMOND
    becu = "/root/reference/test_suite/systems/BeCu_bulk"
    if case == "si8_orth":
        n, lv, orth, files, types, frac = (16, 16, 18), np.diag([A_SI, A_SI, 1.1 * A_SI]), 1, [psp("C.pot")], \
            [0] * 8, DIAMOND
    elif case == "si2_fcc":
        n, lv, orth, files, types, frac = (15, 15, 15), A_SI * np.array([[0, .5, .5], [.5, 0, .5], [.5, .5, 0]]), 0, \
            [psp("C.pot")], [0, 0], np.array([[0, 0, 0], [.25, .25, .25]])
    else:
        n, lv, orth, files, types, frac = (21, 21, 21), np.diag([5.080050872061664] * 3), 1, \
            [os.path.join(becu, "Be.pot"), os.path.join(becu, "Cu.pot")], [0, 1], np.array([[0, 0, 0], [.5, .5, .5]])
        if not os.path.isdir(becu):
            pytest.skip("reference test_suite absent")
    nd = int(np.prod(n))
    lv = np.ascontiguousarray(lv, float)
    fr = np.ascontiguousarray(frac, float)
    ty = np.ascontiguousarray(types, np.int32)
    arr = (ctypes.c_char_p * len(files))(*[f.encode() for f in files])
    out = [np.zeros(nd) for _ in range(5)]
    sc = np.zeros(5)
    assert ref.ref_electrostatics(n[0], n[1], n[2], P(lv), orth, 12, len(files), arr, len(ty), P(ty), P(fr),
                                  *[P(o) for o in out], P(sc)) == 0
    rb, rbr, rvc, rrho, rcore = out
    lat = oracle.Lattice(lv, bool(orth))
    g = oracle.Grid(*n, lat)
    E = osc.Electrostatics(g, [oracle.load_psp8(f) for f in files], list(types), frac)
    scale = np.abs(rb).max()
    assert len(E.boxes) == int(sc[3])
    assert np.abs(E.b - rb).max() < 1e-12 * scale and np.abs(E.b_ref - rbr).max() < 1e-12 * np.abs(rbr).max()
    assert np.abs(E.vc - rvc).max() < 1e-12 * np.abs(rvc).max()
    assert np.abs(E.rho_at - rrho).max() < 1e-12 * np.abs(rrho).max()
    assert abs(E.eself - sc[0]) < 1e-10 * abs(sc[0]) and abs(E.ec - sc[1]) < 1e-10 * abs(sc[1])
    assert abs(E.int_b - sc[2]) < 1e-9
    # the pseudocharge holds the ionic charge (closed form: the flux of
    # -Z/r through the box boundary)
    assert abs(E.int_b + E.nelectron) < 1e-6
    # the library's host set-up on a host-only context
    ctx = ctypes.c_void_p()
    L.check(L.lib().lynx_ctx_create(ctypes.byref(ctx), -1))
    try:
        L.check(L.lib().lynx_set_grid(ctx, n[0], n[1], n[2], L.dptr(lv), 0 if orth else 1, 12), ctx)
        L.check(L.lib().lynx_set_atoms(ctx, len(files), arr, len(ty), L.iptr(ty), L.dptr(fr), 1), ctx)
        lo = [np.zeros(nd) for _ in range(5)]
        ls = np.zeros(6)
        L.check(L.lib().lynx_get_electrostatics(ctx, *[L.dptr(o) for o in lo], L.dptr(ls)), ctx)
        for mine, theirs in zip(lo, out):
            assert np.abs(mine - theirs).max() <= 1e-12 * max(np.abs(theirs).max(), 1e-300)
        assert abs(ls[0] - sc[0]) < 1e-10 * abs(sc[0]) and abs(ls[1] - sc[1]) < 1e-10 * abs(sc[1])
        assert int(ls[5]) == int(sc[3]) and int(ls[3]) == E.nelectron
    finally:
        L.lib().lynx_ctx_destroy(ctx)
