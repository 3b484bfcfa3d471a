"""GPU parity of the HIP path against the CPU oracle (same seeded inputs).

Bars: FP64 results agree with the oracle to 1e-11 relative (stencil, nonlocal,
filter) and 1e-9 (Ritz values of a whole solve, Lanczos bounds); FP32 to 2e-5
relative.  Every call goes through the C ABI of liblynx_chefsi.so.
"""
import ctypes

import numpy as np
import pytest

from systems import DIAMOND, A_SI, System, displace, si8

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module")
def si8_24():
    s = si8(24)
    return s, s.oracle_op(), s.hamiltonian()


def test_h_apply_local_only_real():
    s = si8(20, with_nonlocal=False)
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(5, 1)
    for c in (0.0, -0.37):
        ref = op.apply(X, s.veff, c=c)
        got = H.apply(H.halo, X.T.copy().ravel(), s.veff, c).reshape(5, -1).T
        assert rel(got, ref) < 1e-12, c


def test_h_apply_nonlocal_real(si8_24):
    s, op, H = si8_24
    X = s.random_block(6, 2)
    ref = op.apply(X, s.veff)
    got = H.apply(H.halo, X.T.copy(), s.veff).reshape(6, -1).T
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("n", [(18, 20, 24), (60, 60, 20), (26, 90, 12)])
def test_h_apply_bands_and_shapes(n):
    # shapes that force several row bands and partial last bands
    s = System(n, A_SI, DIAMOND)
    s.latvec = np.diag([A_SI * n[0] / 24, A_SI * n[1] / 24, A_SI * n[2] / 24])
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(3, 3)
    ref = op.apply(X, s.veff, c=0.25)
    got = H.apply(H.halo, X.T.copy(), s.veff, 0.25).reshape(3, -1).T
    assert rel(got, ref) < 1e-12


def test_h_apply_kpoint():
    k = (0.25 * 2 * np.pi / A_SI, -0.125 * 2 * np.pi / A_SI, 0.5 * 2 * np.pi / A_SI)
    s = si8(20, kpt=k)
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(4, 4, complex_=True)
    ref = op.apply(X, s.veff, c=0.1)
    got = H.apply_kpt(H.halo, X.T.copy(), s.veff, k, 0.1).reshape(4, -1).T
    assert rel(got, ref) < 1e-12
    # Hermiticity of the Bloch Hamiltonian
    HX = got
    m = X.conj().T @ HX
    assert np.abs(m - m.conj().T).max() < 1e-9 * np.abs(m).max()


def test_h_apply_fp32(si8_24):
    s, op, H = si8_24
    X = s.random_block(4, 5)
    ref = op.apply(X, s.veff)
    got = H._apply(X.T.copy(), s.veff, 0.0, np.float32).reshape(4, -1).T
    assert rel(got, ref) < 2e-6


def test_filter_real(si8_24):
    s, op, H = si8_24
    from lynx_amd import _lib as L
    X = s.random_block(8, 6)
    lo, hi = -0.4, 45.0
    lc = 0.6
    for deg in (1, 2, 7, 20):
        ref = op.cheb_filter(X, s.veff, lc, lo, hi, deg)
        H.set_veff(s.veff)
        Y = np.zeros_like(X)
        L.check(L.lib().lynx_cheb_filter(H.ctx, L.F64, L.vptr(X), X.shape[0], L.vptr(Y), X.shape[0], 8, deg,
                                         lc, lo, hi, L.MEM_HOST), H.ctx)
        assert rel(Y, ref) < 1e-11, deg


def test_filter_complex_kpoint():
    k = (0.5 * 2 * np.pi / A_SI, 0.0, 0.25 * 2 * np.pi / A_SI)
    s = si8(20, kpt=k)
    op, H = s.oracle_op(), s.hamiltonian()
    from lynx_amd import _lib as L
    X = s.random_block(4, 7, complex_=True)
    ref = op.cheb_filter(X, s.veff, 0.5, -0.5, 40.0, 12)
    H.set_veff(s.veff)
    Y = np.zeros_like(X)
    L.check(L.lib().lynx_cheb_filter(H.ctx, L.C128, L.vptr(X), X.shape[0], L.vptr(Y), X.shape[0], 4, 12,
                                     0.5, -0.5, 40.0, L.MEM_HOST), H.ctx)
    assert rel(Y, ref) < 1e-11


def test_filter_fp32(si8_24):
    s, op, H = si8_24
    from lynx_amd import _lib as L
    X = s.random_block(4, 8)
    ref = op.cheb_filter(X, s.veff, 0.6, -0.4, 45.0, 10)
    H.set_veff(s.veff)
    X32 = np.asfortranarray(X, np.float32)
    Y = np.zeros_like(X32)
    L.check(L.lib().lynx_cheb_filter(H.ctx, L.F32, L.vptr(X32), X.shape[0], L.vptr(Y), X.shape[0], 4, 10,
                                     0.6, -0.4, 45.0, L.MEM_HOST), H.ctx)
    assert rel(Y, ref) < 2e-5


def test_set_veff_refuses_non_finite_values(si8_24):
    """lynx_set_veff checks finiteness on the device: a NaN or infinite entry
    is refused with LYNX_E_NONFINITE and leaves the context without a Veff
    (LYNX_E_STATE until the next good call); a huge finite entry is accepted.
    After a good Veff the FP32 filter, which reads the FP32 copy made on the
    device, matches the oracle again."""
    s, op, H = si8_24
    from lynx_amd import _lib as L
    lib = L.lib()
    X = s.random_block(4, 8)
    X32 = np.asfortranarray(X, np.float32)
    Y = np.zeros_like(X32)

    def run():
        return lib.lynx_cheb_filter(H.ctx, L.F32, L.vptr(X32), X.shape[0], L.vptr(Y), X.shape[0], 4, 10,
                                    0.6, -0.4, 45.0, L.MEM_HOST)

    for pos, bad in ((len(s.veff) // 3, np.nan), (0, np.inf), (len(s.veff) - 1, -np.inf)):
        v = s.veff.copy()
        v[pos] = bad
        assert lib.lynx_set_veff(H.ctx, L.dptr(v), L.MEM_HOST) == L.E_NONFINITE, bad
        assert run() == L.E_STATE, bad
    v = s.veff.copy()
    v[5] = 1e300
    assert lib.lynx_set_veff(H.ctx, L.dptr(v), L.MEM_HOST) == L.OK
    # through the C ABI: the wrapper's set_veff skips an array it uploaded last
    good = np.ascontiguousarray(s.veff, np.float64)
    assert lib.lynx_set_veff(H.ctx, L.dptr(good), L.MEM_HOST) == L.OK
    H._veff_key = None
    L.check(run(), H.ctx)
    ref = op.cheb_filter(X, s.veff, 0.6, -0.4, 45.0, 10)
    assert rel(Y, ref) < 2e-5
    # the wrapper: a refused potential raises, and the array uploaded before
    # it is sent again by the next call
    H.set_veff(s.veff)
    v = s.veff.copy()
    v[7] = np.nan
    with pytest.raises(L.LynxError):
        H.set_veff(v)
    assert run() == L.E_STATE
    H.set_veff(s.veff)
    Y[:] = 0
    L.check(run(), H.ctx)
    assert rel(Y, ref) < 2e-5


def test_set_veff_in_place_change_is_uploaded(si8_24):
    """The wrapper skips an upload only for bit-identical values: a change
    made in place to one entry of the same array reaches the device."""
    s, op, H = si8_24
    X = s.random_block(3, 4)
    v = s.veff.copy()
    H.set_veff(v)
    v[1] += 0.75
    got = H.apply(H.halo, X.T.copy().ravel(), v, 0.0).reshape(3, -1).T
    assert rel(got, op.apply(X, v)) < 1e-11
    H.set_veff(s.veff)


def _subspace_dist(A, B, dV):
    # distance between the spans: || P_A - P_B || via orthonormal bases
    qa, _ = np.linalg.qr(A)
    qb, _ = np.linalg.qr(B)
    s = np.linalg.svd(qa.conj().T @ qb, compute_uv=False)
    return float(np.sqrt(max(0.0, 1.0 - s.min() ** 2)))


def test_solve_real(si8_24):
    s, op, H = si8_24
    from lynx_amd import core
    es = core.EigenSolver().setup(s.context(16), H)
    X = s.random_block(16, 9)
    lo, hi, _ = op.lanczos_bounds(s.veff)
    lc = 0.5 * (lo + hi)
    ref_psi, ref_eig, ref_lc = op.solve(X, s.veff, lc, lo, hi, 12)
    psi, eig = es.solve(X, s.veff, lc, lo, hi, 12)
    assert np.abs(eig - ref_eig).max() < 1e-9 * max(1.0, np.abs(ref_eig).max())
    assert abs(es.lambda_cutoff() - ref_lc) < 1e-9
    # orthonormality in the dV metric and same subspace
    dV = s.a ** 3 / s.nd
    G = dV * psi.T @ psi
    assert np.abs(G - np.eye(16)).max() < 1e-10
    assert _subspace_dist(psi, ref_psi, dV) < 1e-7


def test_solve_repeated_converges(si8_24):
    # repeated CheFSI passes converge; eigenvalues equal the oracle's pass by pass
    s, op, H = si8_24
    from lynx_amd import core
    es = core.EigenSolver().setup(s.context(12), H)
    X = s.random_block(12, 10)
    lo, hi, _ = op.lanczos_bounds(s.veff)
    lc = 0.5 * (lo + hi)
    psi, rpsi = X, X
    lcg = lcr = lc
    for it in range(4):
        psi, eig = es.solve(psi, s.veff, lcg, lo, hi, 15)
        rpsi, reig, lcr = op.solve(rpsi, s.veff, lcr, lo, hi, 15)
        lcg = es.lambda_cutoff()
        assert np.abs(eig - reig).max() < 1e-8 * max(1.0, np.abs(reig).max()), it


def test_solve_kpoint():
    k = (0.25 * 2 * np.pi / A_SI, 0.25 * 2 * np.pi / A_SI, 0.0)
    s = si8(20, kpt=k)
    op, H = s.oracle_op(), s.hamiltonian()
    from lynx_amd import core
    es = core.EigenSolver().setup(s.context(10), H)
    X = s.random_block(10, 11, complex_=True)
    lo, hi, _ = op.lanczos_bounds(s.veff)
    lc = 0.5 * (lo + hi)
    ref_psi, ref_eig, _ = op.solve(X, s.veff, lc, lo, hi, 10)
    psi, eig = es.solve_kpt(X, s.veff, lc, lo, hi, k, cheb_degree=10)
    assert np.abs(eig - ref_eig).max() < 1e-9 * max(1.0, np.abs(ref_eig).max())


def test_lanczos_real(si8_24):
    s, op, H = si8_24
    from lynx_amd import core
    es = core.EigenSolver().setup(s.context(8), H)
    mn, mx = es.lanczos_bounds(s.veff)
    rmn, rmx, rit = op.lanczos_bounds(s.veff)
    assert abs(mn - rmn) < 1e-8 * max(1, abs(rmn))
    assert abs(mx - rmx) < 1e-8 * max(1, abs(rmx))
    assert es.lanczos_iters == rit


def test_lanczos_kpoint():
    k = (0.5 * 2 * np.pi / A_SI, 0.0, 0.0)
    s = si8(20, kpt=k)
    op, H = s.oracle_op(), s.hamiltonian()
    from lynx_amd import core
    es = core.EigenSolver().setup(s.context(8), H)
    mn, mx = es.lanczos_bounds_kpt(s.veff, k)
    rmn, rmx, _ = op.lanczos_bounds(s.veff)
    assert abs(mn - rmn) < 1e-8 * max(1, abs(rmn))
    assert abs(mx - rmx) < 1e-8 * max(1, abs(rmx))


def test_displaced_supercell_apply():
    # Si64-shaped 2x2x2 cell with displaced atoms on a coarse grid
    from systems import supercell
    frac = displace(supercell(2), 2 * A_SI, 0.05, seed=1)
    s = System(32, 2 * A_SI, frac)
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(4, 12)
    ref = op.apply(X, s.veff)
    got = H.apply(H.halo, X.T.copy(), s.veff).reshape(4, -1).T
    assert rel(got, ref) < 1e-12


def test_device_pointers_and_leading_dimension():
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray, synchronize
    s = si8(20)
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(3, 13)
    ref = op.apply(X, s.veff, c=0.2)
    nd = s.nd
    ld = nd + 37
    H.set_veff(s.veff)
    hx = np.zeros((3, ld))
    hx[:, :nd] = X.T
    dx = DeviceArray.from_numpy(hx)
    dy = DeviceArray.from_numpy(np.full((3, ld), 7.0))
    L.check(L.lib().lynx_h_apply(H.ctx, L.F64, dx.value, ld, dy.value, ld, 3, 0.2, L.MEM_DEVICE), H.ctx)
    synchronize()
    out = dy.download(np.zeros((3, ld)))
    assert rel(out[:, :nd].T, ref) < 1e-12
    assert np.all(out[:, nd:] == 7.0)    # padding untouched
    # filter and solve on device pointers with padded leading dimension
    Yd = DeviceArray.from_numpy(np.zeros((3, ld)))
    L.check(L.lib().lynx_cheb_filter(H.ctx, L.F64, dx.value, ld, Yd.value, ld, 3, 9, 0.6, -0.4, 45.0,
                                     L.MEM_DEVICE), H.ctx)
    ref_f = op.cheb_filter(X, s.veff, 0.6, -0.4, 45.0, 9)
    out = Yd.download(np.zeros((3, ld)))
    assert rel(out[:, :nd].T, ref_f) < 1e-11


def test_errors_are_reported():
    from lynx_amd import _lib as L
    s = si8(20)
    H = s.hamiltonian()
    X = s.random_block(2, 14)
    Y = np.zeros_like(X)
    # no Veff yet
    st = L.lib().lynx_h_apply(H.ctx, L.F64, L.vptr(X), s.nd, L.vptr(Y), s.nd, 2, 0.0, L.MEM_HOST)
    assert st == L.E_STATE
    H.set_veff(s.veff)
    # leading dimension too small
    st = L.lib().lynx_h_apply(H.ctx, L.F64, L.vptr(X), s.nd - 1, L.vptr(Y), s.nd, 2, 0.0, L.MEM_HOST)
    assert st == L.E_ARG
    # singular overlap: a zero column gives a zero pivot in any summation
    # order -> Cholesky failure status (a duplicated column is singular only
    # up to the GEMM's rounding)
    Z = np.asfortranarray(np.concatenate([X, np.zeros((s.nd, 1))], axis=1))
    st = L.lib().lynx_orthogonalize(H.ctx, L.F64, L.vptr(Z), s.nd, 3, L.MEM_HOST)
    assert st == L.E_CHOLESKY
    # non-finite potential
    v = s.veff.copy()
    v[5] = np.nan
    st = L.lib().lynx_set_veff(H.ctx, L.dptr(v), L.MEM_HOST)
    assert st == L.E_NONFINITE


@pytest.mark.parametrize("n", [(24, 24, 24), (60, 40, 26), (90, 30, 24), (26, 36, 24)])
def test_every_stencil_launch_shape(n):
    # each launch shape the tuner may pick (points per thread, z chunks,
    # halo slots, LDS pitch) gives the oracle's result, real and complex;
    # nx = 90 and 26 take 8-byte float runs (nx not a multiple of four)
    from lynx_amd import _lib as L
    k = (0.25 * 2 * np.pi / A_SI, 0.0, -0.125 * 2 * np.pi / A_SI)
    s = System(n, A_SI, DIAMOND, kpt=k)
    s.latvec = np.diag([A_SI * n[0] / 24, A_SI * n[1] / 24, A_SI * n[2] / 24])
    op, H = s.oracle_op(), s.hamiltonian()
    Xr = s.random_block(2, 21)
    Xc = s.random_block(2, 22, complex_=True)
    op.set_kpoint((0, 0, 0))
    ref_r = op.apply(Xr, s.veff, c=0.1)
    op.set_kpoint(k)
    ref_c = op.apply(Xc, s.veff, c=0.1)
    shapes = [(v, t, d, z) for v in (0, 1, 2, 3, 5, 6, 7) for t in (256, 512) for d in (1, 2) for z in (1, 2, 3, 4)]
    for shape in shapes:
        L.check(L.lib().lynx_set_stencil_config(H.ctx, *shape[1:], shape[0]), H.ctx)
        H.set_kpoint((0, 0, 0))
        got = H.apply(H.halo, Xr.T.copy(), s.veff, 0.1).reshape(2, -1).T
        assert rel(got, ref_r) < 1e-12, shape
        got = H.apply_kpt(H.halo, Xc.T.copy(), s.veff, k, 0.1).reshape(2, -1).T
        assert rel(got, ref_c) < 1e-12, shape
        g64 = H._apply(Xc.T.copy(), s.veff, 0.1, np.complex64).reshape(2, -1).T
        assert rel(g64, ref_c) < 2e-6, shape
        H.set_kpoint((0, 0, 0))
        g32 = H._apply(Xr.T.copy(), s.veff, 0.1, np.float32).reshape(2, -1).T
        assert rel(g32, ref_r) < 2e-6, shape
    L.check(L.lib().lynx_set_stencil_config(H.ctx, 0, 0, 0, 0), H.ctx)


def test_every_launch_shape_twenty_columns():
    # the single-precision kernels on a block as wide as the eigen-stage
    # tests use (20 columns: several column groups per launch), every shape
    from lynx_amd import _lib as L
    k = (0.25 * 2 * np.pi / A_SI, -0.125 * 2 * np.pi / A_SI, 0.375 * 2 * np.pi / A_SI)
    s = System(16, A_SI, DIAMOND, kpt=k)
    op, H = s.oracle_op(), s.hamiltonian()
    Xr = s.random_block(20, 23)
    Xc = s.random_block(20, 24, complex_=True)
    op.set_kpoint((0, 0, 0))
    ref_r = op.apply(Xr, s.veff, c=0.1)
    op.set_kpoint(k)
    ref_c = op.apply(Xc, s.veff, c=0.1)
    shapes = [(v, t, d, z) for v in (0, 3, 7) for t in (256, 512) for d in (1, 2) for z in (1, 2, 3, 4)]
    for shape in shapes:
        L.check(L.lib().lynx_set_stencil_config(H.ctx, *shape[1:], shape[0]), H.ctx)
        H.set_kpoint(k)
        g64 = H._apply(Xc.T.copy(), s.veff, 0.1, np.complex64).reshape(20, -1).T
        assert rel(g64, ref_c) < 2e-6, shape
        got = H._apply(Xc.T.copy(), s.veff, 0.1, np.complex128).reshape(20, -1).T
        assert rel(got, ref_c) < 1e-12, shape
        H.set_kpoint((0, 0, 0))
        g32 = H._apply(Xr.T.copy(), s.veff, 0.1, np.float32).reshape(20, -1).T
        assert rel(g32, ref_r) < 2e-6, shape
    L.check(L.lib().lynx_set_stencil_config(H.ctx, 0, 0, 0, 0), H.ctx)


def test_multispecies_large_spheres_apply():
    # BaTiO3 stand-in (La, Zr, O psp8): 32 projectors per atom and spheres of
    # several thousand points, so the gather splits each atom over four waves
    from systems import batio3
    s = batio3(0, n=28)
    op, H = s.oracle_op(), s.hamiltonian()
    assert max(len(im.gpos) for im in op.images) > 1024
    X = s.random_block(5, 41)
    ref = op.apply(X, s.veff, c=0.05)
    got = H.apply(H.halo, X.T.copy(), s.veff, 0.05).reshape(5, -1).T
    assert rel(got, ref) < 1e-12
    k = (0.5 * 2 * np.pi / s.a, 0.0, 0.25 * 2 * np.pi / s.a)
    op.set_kpoint(k)
    Xc = s.random_block(3, 42, complex_=True)
    refc = op.apply(Xc, s.veff)
    gotc = H.apply_kpt(H.halo, Xc.T.copy(), s.veff, k).reshape(3, -1).T
    assert rel(gotc, refc) < 1e-12
    g32 = H._apply(Xc.T.copy(), s.veff, 0.0, np.complex64).reshape(3, -1).T
    assert rel(g32, refc) < 2e-6


def test_wide_blocks_nonlocal():
    # 36-40 columns (ten column groups of the gather and scatter), real and
    # complex, FP64 and single precision, 8 projectors (Si8) and 32 with
    # spheres of thousands of points (BaTiO3 stand-in)
    from systems import batio3
    for s, ncol in ((si8(16), 40), (batio3(0, n=28), 36)):
        k = (0.25 * 2 * np.pi / s.a, -0.125 * 2 * np.pi / s.a, 0.375 * 2 * np.pi / s.a)
        op, H = s.oracle_op(), s.hamiltonian()
        X = s.random_block(ncol, 43)
        ref = op.apply(X, s.veff, c=0.05)
        got = H.apply(H.halo, X.T.copy(), s.veff, 0.05).reshape(ncol, -1).T
        assert rel(got, ref) < 1e-12
        g32 = H._apply(X.T.copy(), s.veff, 0.05, np.float32).reshape(ncol, -1).T
        assert rel(g32, ref) < 2e-6
        op.set_kpoint(k)
        Xc = s.random_block(ncol, 44, complex_=True)
        refc = op.apply(Xc, s.veff)
        gotc = H.apply_kpt(H.halo, Xc.T.copy(), s.veff, k).reshape(ncol, -1).T
        assert rel(gotc, refc) < 1e-12
        g64 = H._apply(Xc.T.copy(), s.veff, 0.0, np.complex64).reshape(ncol, -1).T
        assert rel(g64, refc) < 2e-6


def test_filter_graph_replays_match_direct_launches():
    # the second filter call with the same buffers captures a HIP graph and
    # later calls replay it with their own (a, b, shift) table: every call,
    # with other GPU work and synchronisations between them, against the
    # oracle, and bit-equal to direct launches of the same call; then the
    # same with HIP-event timing on (its own graph with event records)
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray, synchronize
    s = si8(20)
    op, H = s.oracle_op(), s.hamiltonian()
    H.set_veff(s.veff)
    lib = L.lib()
    k = (0.25 * 2 * np.pi / A_SI, 0.0, -0.125 * 2 * np.pi / A_SI)
    cases = ((np.float64, L.F64, 1e-11, None), (np.float32, L.F32, 2e-5, None),
             (np.complex128, L.C128, 1e-11, k), (np.complex64, L.C64, 2e-5, k))
    for dt, code, tol, kp in cases:
        H.set_kpoint(kp if kp is not None else (0.0, 0.0, 0.0))
        op.set_kpoint(kp if kp is not None else (0.0, 0.0, 0.0))
        X = s.random_block(6, 51, complex_=kp is not None)
        xd = DeviceArray.from_numpy(np.ascontiguousarray(X.T.astype(dt)))
        yd = DeviceArray(X.size * np.dtype(dt).itemsize)
        params = [(0.30, -0.6, 40.0, 9), (0.45, -0.5, 42.0, 9), (0.35, -0.7, 39.0, 9), (0.30, -0.6, 40.0, 9)]
        for i, (lc, lo, hi, deg) in enumerate(params):
            L.check(lib.lynx_cheb_filter(H.ctx, code, xd.value, s.nd, yd.value, s.nd, 6, deg, lc, lo, hi,
                                         L.MEM_DEVICE), H.ctx)
            got = np.zeros((6, s.nd), dt)
            yd.download(got)
            ref = op.cheb_filter(X, s.veff, lc, lo, hi, deg)
            assert rel(got.T, ref) < tol, (dt, i)
            # direct launches of the same call
            L.check(lib.lynx_set_filter_graphs(H.ctx, 0), H.ctx)
            L.check(lib.lynx_cheb_filter(H.ctx, code, xd.value, s.nd, yd.value, s.nd, 6, deg, lc, lo, hi,
                                         L.MEM_DEVICE), H.ctx)
            L.check(lib.lynx_set_filter_graphs(H.ctx, 1), H.ctx)
            direct = np.zeros((6, s.nd), dt)
            yd.download(direct)
            assert np.array_equal(got, direct), (dt, i)
            # unrelated GPU work between the calls
            H._apply(X.T.copy(), s.veff, 0.0, dt)
            synchronize()
        L.check(lib.lynx_set_timing(H.ctx, 1), H.ctx)
        t_d = (ctypes.c_double * 4)()
        t_n = (ctypes.c_int * 4)()
        for i in range(3):
            lc, lo, hi, deg = params[i]
            L.check(lib.lynx_cheb_filter(H.ctx, code, xd.value, s.nd, yd.value, s.nd, 6, deg, lc, lo, hi,
                                         L.MEM_DEVICE), H.ctx)
            L.check(lib.lynx_last_filter_timing(H.ctx, t_d, t_n), H.ctx)
            assert t_n[1] == deg and t_d[1] > 0, (dt, i)
            got = np.zeros((6, s.nd), dt)
            yd.download(got)
            assert rel(got.T, op.cheb_filter(X, s.veff, lc, lo, hi, deg)) < tol, (dt, "timed", i)
        L.check(lib.lynx_set_timing(H.ctx, 0), H.ctx)
        xd.free()
        yd.free()


@pytest.mark.parametrize("n", [(10, 9, 8), (14, 7, 13)])
def test_small_grids_wrap_within_the_stencil_radius(n):
    # grids no wider than 2R = 12 points along some axis: the halo rows and
    # planes wrap onto the block's own rows (R = 6 < every dimension)
    s = System(n, A_SI, DIAMOND, kpt=(0.3, -0.2, 0.1))
    s.latvec = np.diag([A_SI * n[0] / 16, A_SI * n[1] / 16, A_SI * n[2] / 16])
    op, H = s.oracle_op(), s.hamiltonian()
    X = s.random_block(3, 61)
    op.set_kpoint((0, 0, 0))
    ref = op.apply(X, s.veff, c=0.2)
    H.set_kpoint((0, 0, 0))
    got = H.apply(H.halo, X.T.copy(), s.veff, 0.2).reshape(3, -1).T
    assert rel(got, ref) < 1e-12
    k = (0.3, -0.2, 0.1)
    op.set_kpoint(k)
    Xc = s.random_block(3, 62, complex_=True)
    refc = op.apply(Xc, s.veff)
    gotc = H.apply_kpt(H.halo, Xc.T.copy(), s.veff, k).reshape(3, -1).T
    assert rel(gotc, refc) < 1e-12
