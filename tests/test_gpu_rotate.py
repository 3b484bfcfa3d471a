"""Rotations through lynx_rotate (c = a q, the rotation step of the CheFSI
pass) against numpy in FP64 on the same inputs.  Errors per entry are
measured against the entry's scale |a_i| |q_j| (row i of a, column j of q):

- FP64 blocks: rocBLAS GEMM, 1e-14;
- FP32 blocks with LYNX_GRAM_FP32: q rounded to FP32, FP32 GEMM, 5e-7;
- FP32 blocks with LYNX_GRAM_BF16X3_ROT (the hand-written split-BF16x3
  MFMA kernel, q split from FP64): 2e-5 (each operand is held to about 2^-17,
  the products accumulate in FP32); exactly bfloat16 a and q (the tails
  vanish) to the FP32 accumulation alone, 5e-7;
- shapes that exercise the kernel's edges: one, two and three 160-column
  strips, a ragged last 32-column block, more columns than one workgroup
  row holds (n > 480), grid sizes that are not a multiple of the 128-point
  workgroup tile, host and device operands.
"""
import ctypes

import numpy as np
import pytest

from systems import System

pytestmark = pytest.mark.gpu


def bf16_round(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)


def _rotate(L, H, dtype, A, Q, mem="host"):
    nd, n = A.shape
    A = np.asfortranarray(A)
    Qf = np.asfortranarray(Q, np.float64)
    C = np.zeros_like(A)
    if mem == "host":
        L.check(L.lib().lynx_rotate(H.ctx, dtype, A.ctypes.data_as(ctypes.c_void_p), nd, L.dptr(Qf), n,
                                    C.ctypes.data_as(ctypes.c_void_p), nd, L.MEM_HOST), H.ctx)
        return C
    from lynx_amd.hipmem import DeviceArray, synchronize
    dA = DeviceArray.from_numpy(A.T.copy())        # one column per row of the copy: column-major on the device
    dC = DeviceArray.from_numpy(np.zeros((n, nd), A.dtype))
    L.check(L.lib().lynx_rotate(H.ctx, dtype, dA.value, nd, L.dptr(Qf), n, dC.value, nd, L.MEM_DEVICE), H.ctx)
    synchronize()
    out = dC.download(np.zeros((n, nd), A.dtype)).T
    dA.free()
    dC.free()
    return out


@pytest.mark.parametrize("grid,n", [((16, 16, 16), 24), ((20, 21, 22), 200), ((30, 30, 31), 480),
                                    ((24, 25, 26), 161), ((18, 18, 19), 520)])
def test_rotation_matches_fp64(grid, n):
    from lynx_amd import _lib as L
    s = System(grid, 10.0, np.zeros((0, 3)), with_nonlocal=False)
    H = s.hamiltonian(0)
    nd = s.nd
    rng = np.random.default_rng(n + nd)
    A = rng.uniform(-0.5, 0.5, (nd, n))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    Q = Q * rng.uniform(0.5, 2.0, n)               # not orthogonal: columns of different size
    scale = np.abs(A).sum(axis=1)[:, None] * np.abs(Q).max(axis=0)[None, :]

    def err(got, a, q):
        return float((np.abs(got.astype(np.float64) - a.astype(np.float64) @ q) / scale).max())

    e64 = err(_rotate(L, H, L.F64, A, Q), A, Q)
    A32 = A.astype(np.float32)
    out = {}
    for mode in (L.GRAM_FP32, L.GRAM_BF16X3, L.GRAM_BF16X3_ROT):
        L.check(L.lib().lynx_set_gram_precision(H.ctx, mode), H.ctx)
        out[mode] = _rotate(L, H, L.F32, A32, Q)
    e_fp32 = err(out[L.GRAM_FP32], A32, Q)
    e_bf3 = err(out[L.GRAM_BF16X3_ROT], A32, Q)
    # exactly bfloat16 operands: only the FP32 accumulation rounds
    Ab, Qb = bf16_round(A32), bf16_round(Q).astype(np.float64)
    e_exact = err(_rotate(L, H, L.F32, Ab, Qb), Ab, Qb)
    dev = _rotate(L, H, L.F32, A32, Q, mem="device")
    print(f"[rotate] grid {grid} n {n}: fp64 {e64:.2e} fp32 {e_fp32:.2e} bf16x3 {e_bf3:.2e} exact-bf16 {e_exact:.2e}",
          flush=True)
    assert e64 < 1e-14
    assert e_fp32 < 5e-7
    assert e_bf3 < 2e-5
    assert e_exact < 5e-7
    assert np.array_equal(dev, out[L.GRAM_BF16X3_ROT])
    # LYNX_GRAM_BF16X3 leaves the rotation an FP32 GEMM; the split changes it
    assert np.array_equal(out[L.GRAM_FP32], out[L.GRAM_BF16X3])
    assert not np.array_equal(out[L.GRAM_FP32], out[L.GRAM_BF16X3_ROT])
    L.check(L.lib().lynx_set_gram_precision(H.ctx, L.GRAM_FP32), H.ctx)
