"""lynx_h_apply_noalloc (the H apply that allocates nothing: its projector
coefficients live in the library's per-GPU scratch array) against
lynx_h_apply on the same device buffers: two calls in a row, the second with
fewer columns, for real FP64 and FP32 blocks at Gamma and complex128 blocks
at a k-point.  With the launch shape both calls share, the results are equal
bit for bit; both are also checked against the oracle (1e-12, FP32 2e-6).
A column count whose coefficients exceed the scratch array is refused before
any launch."""
import ctypes

import numpy as np
import pytest

from systems import A_SI, si8

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("case", ["f64", "f32", "c128"])
def test_noalloc_apply_matches_apply(case):
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray
    lib = L.lib()
    kpt = (0.25 * 2 * np.pi / A_SI, -0.125 * 2 * np.pi / A_SI, 0.5 * 2 * np.pi / A_SI) if case == "c128" else (0, 0, 0)
    s = si8(24, kpt=kpt)
    H = s.hamiltonian()
    H.set_veff(s.veff)
    op = s.oracle_op()
    code = {"f64": L.F64, "f32": L.F32, "c128": L.C128}[case]
    dt = L.np_dtype(code)
    nd = s.nd
    for ncol, c in ((6, 0.0), (3, -0.25)):
        X = s.random_block(ncol, 40 + ncol, complex_=(case == "c128"))
        Xd = DeviceArray.from_numpy(np.ascontiguousarray(X.T, dtype=dt))
        nbytes = X.size * np.dtype(dt).itemsize
        Y1, Y2 = DeviceArray(nbytes), DeviceArray(nbytes)
        print(f"[no-alloc] {case} ncol={ncol}", flush=True)
        L.check(lib.lynx_h_apply(H.ctx, code, Xd.value, nd, Y1.value, nd, ncol, c, L.MEM_DEVICE), H.ctx)
        L.check(lib.lynx_h_apply_noalloc(H.ctx, code, Xd.value, nd, Y2.value, nd, ncol, c), H.ctx)
        L.check(lib.lynx_ctx_synchronize(H.ctx), H.ctx)
        a = np.zeros((ncol, nd), dt)
        b = np.zeros((ncol, nd), dt)
        Y1.download(a)
        Y2.download(b)
        assert np.array_equal(a, b), (case, ncol)
        ref = op.apply(X, s.veff, c=c)
        assert rel(b.T, ref) < (2e-6 if case == "f32" else 1e-12), (case, ncol)
        for d in (Xd, Y1, Y2):
            d.free()
    # coefficients of too many columns for the scratch array: refused on the
    # host, before anything is launched (the buffers are never touched)
    n_atom, npm, es = 8, 8, np.dtype(dt).itemsize
    too_many = (32 << 20) // (n_atom * npm * es) + 1
    dummy = DeviceArray(64)
    st = lib.lynx_h_apply_noalloc(H.ctx, code, dummy.value, nd, dummy.value, nd, too_many, 0.0)
    assert st == L.E_ARG
    assert b"scratch" in lib.lynx_last_error(H.ctx)
    dummy.free()
