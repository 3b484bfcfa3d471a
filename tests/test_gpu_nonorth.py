"""GPU parity for non-orthogonal cells (FCC primitive Si2): the fused stencil
with the three mixed-derivative terms, with and without Bloch phases, against
the oracle's direct evaluation of Hamiltonian.cpp:176-267.  Same bars as the
orthogonal path: FP64 1e-12 per apply and 1e-11 per filter, FP32 2e-6 per
apply."""
import numpy as np
import pytest

from systems import si2_fcc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def kvec(s, red):
    L = np.sqrt((s.latvec ** 2).sum(axis=1))
    return tuple(2 * np.pi * np.asarray(red) / L)


@pytest.mark.parametrize("n", [24, (20, 24, 28), (30, 18, 16)])
def test_nonorth_apply_every_launch_shape(n):
    from lynx_amd import _lib as L
    s = si2_fcc(n)
    k = kvec(s, (0.25, -0.125, 0.375))
    s.kpt = k
    op, H = s.oracle_op(), s.hamiltonian()
    Xr = s.random_block(3, 31)
    Xc = s.random_block(3, 32, complex_=True)
    op.set_kpoint((0, 0, 0))
    ref_r = op.apply(Xr, s.veff, c=-0.2)
    op.set_kpoint(k)
    ref_c = op.apply(Xc, s.veff, c=-0.2)
    shapes = [(t, d, z) for t in (256, 512) for d in (1, 2) for z in (1, 2)]
    for shape in shapes:
        L.check(L.lib().lynx_set_stencil_config(H.ctx, *shape, 0), H.ctx)
        H.set_kpoint((0, 0, 0))
        got = H.apply(H.halo, Xr.T.copy(), s.veff, -0.2).reshape(3, -1).T
        assert rel(got, ref_r) < 1e-12, shape
        g32 = H._apply(Xr.T.copy(), s.veff, -0.2, np.float32).reshape(3, -1).T
        assert rel(g32, ref_r) < 2e-6, shape
        got = H.apply_kpt(H.halo, Xc.T.copy(), s.veff, k, -0.2).reshape(3, -1).T
        assert rel(got, ref_c) < 1e-12, shape
        g64 = H._apply(Xc.T.copy(), s.veff, -0.2, np.complex64).reshape(3, -1).T
        assert rel(g64, ref_c) < 2e-6, shape
    L.check(L.lib().lynx_set_stencil_config(H.ctx, 0, 0, 0, 0), H.ctx)


@pytest.mark.parametrize("kred", [(0.0, 0.0, 0.0), (0.25, 0.25, 0.0)])
def test_nonorth_single_column_apply(kred):
    # one column, as Lanczos applies H, under the tuned launch shape
    s = si2_fcc(24)
    k = kvec(s, kred)
    s.kpt = k
    op, H = s.oracle_op(), s.hamiltonian()
    cplx = any(abs(v) > 0 for v in k)
    X = s.random_block(1, 37, complex_=cplx)
    ref = op.apply(X, s.veff)
    if cplx:
        got = H.apply_kpt(H.halo, X.T.copy(), s.veff, k).reshape(1, -1).T
    else:
        got = H.apply(H.halo, X.T.copy(), s.veff).reshape(1, -1).T
    assert rel(got, ref) < 1e-12


def test_nonorth_hamiltonian_is_hermitian():
    s = si2_fcc(20)
    k = kvec(s, (0.125, 0.25, -0.375))
    s.kpt = k
    H = s.hamiltonian()
    X = s.random_block(6, 33, complex_=True)
    HX = H.apply_kpt(H.halo, X.T.copy(), s.veff, k, 0.0).reshape(6, -1).T
    m = X.conj().T @ HX
    assert np.abs(m - m.conj().T).max() < 1e-9 * np.abs(m).max()


def test_nonorth_filter_complex_and_c64():
    from lynx_amd import _lib as L
    s = si2_fcc(24)
    k = kvec(s, (0.25, 0.0, 0.125))
    s.kpt = k
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(5, 34, complex_=True)
    lo, hi, _ = op.lanczos_bounds(s.veff)
    lc = 0.5 * (lo + hi)
    ref = op.cheb_filter(X, s.veff, lc, lo, hi, 10)
    H.set_veff(s.veff)
    Y = np.zeros_like(X)
    L.check(L.lib().lynx_cheb_filter(H.ctx, L.C128, L.vptr(X), X.shape[0], L.vptr(Y), X.shape[0], 5, 10,
                                     lc, lo, hi, L.MEM_HOST), H.ctx)
    assert rel(Y, ref) < 1e-11
    X64 = np.asfortranarray(X, np.complex64)
    Y64 = np.zeros_like(X64)
    L.check(L.lib().lynx_cheb_filter(H.ctx, L.C64, L.vptr(X64), X.shape[0], L.vptr(Y64), X.shape[0], 5, 10,
                                     lc, lo, hi, L.MEM_HOST), H.ctx)
    assert rel(Y64, ref) < 2e-5


def test_nonorth_lanczos_and_solve_kpoint():
    from lynx_amd import core
    s = si2_fcc(24)
    k = kvec(s, (0.25, 0.25, 0.0))
    s.kpt = k
    op, H = s.oracle_op(), s.hamiltonian()
    es = core.EigenSolver().setup(s.context(12), H)
    mn, mx = es.lanczos_bounds_kpt(s.veff, k)
    rmn, rmx, rit = op.lanczos_bounds(s.veff)
    print("lanczos", mn, rmn, mx, rmx, es.lanczos_iters, rit, flush=True)
    assert es.lanczos_iters == rit
    assert abs(mn - rmn) < 1e-8 * max(1, abs(rmn))
    assert abs(mx - rmx) < 1e-8 * max(1, abs(rmx))
    X = s.random_block(12, 35, complex_=True)
    lc = 0.5 * (rmn + rmx)
    ref_psi, ref_eig, _ = op.solve(X, s.veff, lc, rmn, rmx, 10)
    psi, eig = es.solve_kpt(X, s.veff, lc, rmn, rmx, k, cheb_degree=10)
    assert np.abs(eig - ref_eig).max() < 1e-9 * max(1.0, np.abs(ref_eig).max())


def test_nonorth_solve_gamma():
    from lynx_amd import core
    s = si2_fcc((20, 24, 28))
    op, H = s.oracle_op(), s.hamiltonian()
    es = core.EigenSolver().setup(s.context(10), H)
    X = s.random_block(10, 36)
    lo, hi, _ = op.lanczos_bounds(s.veff)
    lc = 0.5 * (lo + hi)
    ref_psi, ref_eig, _ = op.solve(X, s.veff, lc, lo, hi, 12)
    psi, eig = es.solve(X, s.veff, lc, lo, hi, cheb_degree=12)
    assert np.abs(eig - ref_eig).max() < 1e-9 * max(1.0, np.abs(ref_eig).max())
