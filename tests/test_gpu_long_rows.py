"""Grids whose x rows are longer than the 64- and 128-point LDS row pitches
hold (nx + 12 > 128): the 512-thread stencil shapes with a 256-point pitch,
against the oracle.

- orthogonal cells with 120- and 200-point rows: every 512-thread shape
  (points per thread, z chunks) and the tuned shape, real and Bloch-phased
  complex blocks, FP64 to 1e-12 and FP32 / complex64 to 2e-6 per apply; the
  degree-20 filter to 1e-11 (FP64);
- a non-orthogonal cell (the FCC primitive cell stretched 5x along its first
  vector, 120 x 24 x 24): real FP64, FP32 and complex64 against the oracle;
  complex128 planes of that pitch do not fit the LDS, and the library
  refuses them with LYNX_E_UNSUPPORTED;
- 256-thread shapes forced on long rows are refused the same way;
- batches (lynx_batch_*) holding long-row units (512-thread groups with a
  256-point pitch) beside short-row ones: every unit's filter against the
  oracle's filter of that unit alone, FP64 / complex128 to 1e-11 and FP32 /
  complex64 to 2e-5; complex128 non-orthogonal long rows refused.
"""
import numpy as np
import pytest

from systems import A_SI, DIAMOND, SI2, FCC_PRIMITIVE, System

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def kvec(latvec, red):
    L = np.sqrt((np.asarray(latvec) ** 2).sum(axis=1))
    return tuple(2 * np.pi * np.asarray(red) / L)


@pytest.mark.parametrize("n", [(120, 24, 20), (200, 16, 16)])
def test_long_rows_orthogonal(n):
    from lynx_amd import _lib as L
    latvec = np.diag([A_SI * n[0] / 24, A_SI * n[1] / 24, A_SI * n[2] / 24])
    k = kvec(latvec, (0.25, -0.125, 0.375))
    s = System(n, A_SI, DIAMOND, kpt=k, latvec=None)
    s.latvec = latvec
    op, H = s.oracle_op(), s.hamiltonian()
    Xr = s.random_block(3, 41)
    Xc = s.random_block(3, 42, complex_=True)
    op.set_kpoint((0, 0, 0))
    ref_r = op.apply(Xr, s.veff, c=0.1)
    op.set_kpoint(k)
    ref_c = op.apply(Xc, s.veff, c=0.1)
    shapes = [(512, d, z) for d in (1, 2) for z in (1, 2, 3)] + [(0, 0, 0)]
    for shape in shapes:
        L.check(L.lib().lynx_set_stencil_config(H.ctx, *shape, 0), H.ctx)
        H.set_kpoint((0, 0, 0))
        got = H.apply(H.halo, Xr.T.copy(), s.veff, 0.1).reshape(3, -1).T
        assert rel(got, ref_r) < 1e-12, shape
        g32 = H._apply(Xr.T.copy(), s.veff, 0.1, np.float32).reshape(3, -1).T
        assert rel(g32, ref_r) < 2e-6, shape
        got = H.apply_kpt(H.halo, Xc.T.copy(), s.veff, k, 0.1).reshape(3, -1).T
        assert rel(got, ref_c) < 1e-12, shape
        g64 = H._apply(Xc.T.copy(), s.veff, 0.1, np.complex64).reshape(3, -1).T
        assert rel(g64, ref_c) < 2e-6, shape
    # a 256-thread shape cannot hold these rows
    L.check(L.lib().lynx_set_stencil_config(H.ctx, 256, 1, 1, 0), H.ctx)
    H.set_kpoint((0, 0, 0))
    y = np.zeros_like(Xr)
    st = L.lib().lynx_h_apply(H.ctx, L.F64, L.vptr(np.asfortranarray(Xr)), s.nd, L.vptr(y), s.nd, 3, 0.1,
                              L.MEM_HOST)
    assert st == L.E_UNSUPPORTED
    L.check(L.lib().lynx_set_stencil_config(H.ctx, 0, 0, 0, 0), H.ctx)
    # the filter, tuned shape
    X = s.random_block(4, 43)
    lo, hi, lc = -0.5, 30.0, 0.8
    op.set_kpoint((0, 0, 0))
    ref = op.cheb_filter(X, s.veff, lc, lo, hi, 20)
    Y = np.zeros_like(X)
    H.set_veff(s.veff)
    for _ in range(3):   # direct launches, graph capture, replay
        L.check(L.lib().lynx_cheb_filter(H.ctx, L.F64, L.vptr(np.asfortranarray(X)), s.nd, L.vptr(Y), s.nd, 4, 20,
                                         lc, lo, hi, L.MEM_HOST), H.ctx)
        assert rel(Y, ref) < 1e-11


def test_long_rows_nonorthogonal():
    from lynx_amd import _lib as L
    lv = A_SI * FCC_PRIMITIVE.copy()
    lv[0] *= 5.0
    n = (120, 24, 24)
    k = kvec(lv, (0.25, 0.125, -0.25))
    s = System(n, A_SI, SI2, latvec=lv, kpt=k)
    op, H = s.oracle_op(), s.hamiltonian()
    Xr = s.random_block(3, 44)
    Xc = s.random_block(3, 45, complex_=True)
    op.set_kpoint((0, 0, 0))
    ref_r = op.apply(Xr, s.veff, c=-0.2)
    op.set_kpoint(k)
    ref_c = op.apply(Xc, s.veff, c=-0.2)
    for shape in [(512, d, z) for d in (1, 2) for z in (1, 2)] + [(0, 0, 0)]:
        L.check(L.lib().lynx_set_stencil_config(H.ctx, *shape, 0), H.ctx)
        H.set_kpoint((0, 0, 0))
        got = H.apply(H.halo, Xr.T.copy(), s.veff, -0.2).reshape(3, -1).T
        assert rel(got, ref_r) < 1e-12, shape
        g32 = H._apply(Xr.T.copy(), s.veff, -0.2, np.float32).reshape(3, -1).T
        assert rel(g32, ref_r) < 2e-6, shape
        H.set_kpoint(k)
        g64 = H._apply(Xc.T.copy(), s.veff, -0.2, np.complex64).reshape(3, -1).T
        assert rel(g64, ref_c) < 2e-6, shape
        y = np.zeros_like(Xc)
        st = L.lib().lynx_h_apply(H.ctx, L.C128, L.vptr(np.asfortranarray(Xc)), s.nd, L.vptr(y), s.nd, 3, -0.2,
                                  L.MEM_HOST)
        assert st == L.E_UNSUPPORTED, shape
    L.check(L.lib().lynx_set_stencil_config(H.ctx, 0, 0, 0, 0), H.ctx)


def _bounds(i):
    return 0.3 + 0.05 * i, -0.6 - 0.02 * i, 60.0 + 3.0 * i


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 2e-5)])
def test_long_rows_batch_real(dtype, tol):
    # 120- and 200-point orthogonal rows, a 120-point non-orthogonal cell and
    # a short-row unit in one batch
    from lynx_amd import core
    lv = A_SI * FCC_PRIMITIVE.copy()
    lv[0] *= 5.0
    systems = [System((120, 24, 20), A_SI, DIAMOND), System((200, 16, 16), A_SI, DIAMOND),
               System((120, 24, 24), A_SI, SI2, latvec=lv), System((20, 20, 20), A_SI, DIAMOND)]
    for s in systems[:2]:
        s.latvec = np.diag([A_SI * s.n[0] / 24, A_SI * s.n[1] / 24, A_SI * s.n[2] / 24])
    Hs = []
    for s in systems:
        H = s.hamiltonian()
        H.set_veff(s.veff)
        Hs.append(H)
    bf = core.BatchFilter([(H, None) for H in Hs])
    blocks = [s.random_block(2 + i % 2, 90 + i) for i, s in enumerate(systems)]
    deg = 7
    lc, lo, hi = zip(*[_bounds(i) for i in range(len(systems))])
    got = bf.filter(blocks, deg, lc, lo, hi, dtype=dtype)
    for i, s in enumerate(systems):
        ref = s.oracle_op().cheb_filter(blocks[i], s.veff, lc[i], lo[i], hi[i], deg)
        assert rel(got[i], ref) < tol, (i, s.n)
    bf.close()


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_long_rows_batch_kpoints(dtype, tol):
    # Bloch-phased k-points of an orthogonal cell with 120-point rows and,
    # in complex64, of the stretched non-orthogonal cell
    from lynx_amd import core, _lib as L
    n = (120, 24, 20)
    latvec = np.diag([A_SI * n[0] / 24, A_SI * n[1] / 24, A_SI * n[2] / 24])
    s = System(n, A_SI, DIAMOND)
    s.latvec = latvec
    H = s.hamiltonian()
    H.set_veff(s.veff)
    units = [(s, H, kvec(latvec, r)) for r in ((0.0, 0.0, 0.0), (0.25, -0.125, 0.375))]
    lv = A_SI * FCC_PRIMITIVE.copy()
    lv[0] *= 5.0
    sn = System((120, 24, 24), A_SI, SI2, latvec=lv)
    Hn = sn.hamiltonian()
    Hn.set_veff(sn.veff)
    kn = kvec(lv, (0.25, 0.125, -0.25))
    if dtype == np.complex64:
        units.append((sn, Hn, kn))
    else:
        with pytest.raises(L.LynxError):
            bad = core.BatchFilter([(Hn, kn)])
            try:
                bad.filter([sn.random_block(2, 99, complex_=True)], 5, [0.3], [-0.6], [60.0], dtype=dtype)
            finally:
                bad.close()
    bf = core.BatchFilter([(h, k) for _, h, k in units])
    blocks = [u[0].random_block(3, 100 + i, complex_=True) for i, u in enumerate(units)]
    deg = 7
    lc, lo, hi = zip(*[_bounds(i) for i in range(len(units))])
    got = bf.filter(blocks, deg, lc, lo, hi, dtype=dtype)
    for i, (sy, _, k) in enumerate(units):
        op = sy.oracle_op()
        op.set_kpoint(k)
        ref = op.cheb_filter(blocks[i], sy.veff, lc[i], lo[i], hi[i], deg)
        assert rel(got[i], ref) < tol, (i, k)
    bf.close()
