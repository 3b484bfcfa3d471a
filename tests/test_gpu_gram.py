"""Gram products through lynx_gram (the overlap / projection product of the
CheFSI pass) against numpy in FP64 on the same inputs:

- FP64 blocks: the overlap A^T A through the hand-written FP64 MFMA kernel
  (upper tiles, mirrored), a general A^T B through rocBLAS split-K GEMMs,
  1e-13 relative; one tile, several 128-column tiles with a ragged last
  one, K not a multiple of the 32-row block, host and device inputs;
- FP32 blocks, errors per entry relative to dV |a_i| |b_j| (the scale of
  a product of two columns): LYNX_GRAM_FP32 (FP32 GEMMs, FP64 chunk sums)
  5e-6 (1.2e-6 measured on a one-chunk sum of squares); LYNX_GRAM_BF16X3
  (hand-written split-BF16x3 MFMA kernel) 2e-5 (the split represents each
  operand to about 2^-17; 2.4e-6 to 4e-6 measured); exactly bfloat16
  inputs (the tail vanishes, only the FP32 accumulation rounds) to 2e-6 of
  the largest entry;
- shapes that exercise the kernel's edges: one partial 160-column tile,
  several tiles with a ragged last one, K not a multiple of the 32-row
  block, the symmetric A^T A and a general A^T B, host and device inputs.
"""
import ctypes

import numpy as np
import pytest

from systems import System

pytestmark = pytest.mark.gpu


def _ctx(n):
    s = System(n, 10.0 * n / 16, np.zeros((0, 3)), with_nonlocal=False)
    H = s.hamiltonian(0)
    return s, H


def _host(x):
    x = np.asfortranarray(x)
    return (x.ctypes.data_as(ctypes.c_void_p), x.shape[0], x)


def bf16_round(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("grid,n,sym", [((16, 16, 16), 20, True), ((20, 21, 22), 200, False),
                                        ((50, 50, 50), 480, True), ((30, 30, 31), 161, False)])
def test_split_bf16x3_gram_matches_fp64(grid, n, sym):
    from lynx_amd import _lib as L
    s = System(grid, 10.0, np.zeros((0, 3)), with_nonlocal=False)
    H = s.hamiltonian(0)
    nd = s.nd
    rng = np.random.default_rng(n)
    A = rng.uniform(-0.5, 0.5, (nd, n)).astype(np.float32)
    B = A if sym else rng.uniform(-0.5, 0.5, (nd, n)).astype(np.float32)
    dV = s.latvec[0, 0] * s.latvec[1, 1] * s.latvec[2, 2] / nd
    ref = dV * (A.astype(np.float64).T @ B.astype(np.float64))
    # per-entry error scale of a product of columns: dV |a_i| |b_j|
    scale = dV * np.linalg.norm(A.astype(np.float64), axis=0).max() * np.linalg.norm(B.astype(np.float64), axis=0).max()
    ha = _host(A)
    hb = ha if sym else _host(B)
    out = {}
    for mode in (L.GRAM_FP32, L.GRAM_BF16X3):
        L.check(L.lib().lynx_set_gram_precision(H.ctx, mode), H.ctx)
        c = np.zeros((n, n))
        L.check(L.lib().lynx_gram(H.ctx, L.F32, ha[0], ha[1], hb[0], hb[1], n, L.dptr(c), L.MEM_HOST), H.ctx)
        out[mode] = np.abs(c.T - ref).max() / scale
    print(f"[gram] {grid} n={n}: FP32 GEMM {out[L.GRAM_FP32]:.1e}, split-BF16x3 {out[L.GRAM_BF16X3]:.1e}", flush=True)
    assert out[L.GRAM_FP32] < 5e-6
    assert out[L.GRAM_BF16X3] < 2e-5


def test_split_bf16x3_gram_exact_on_bfloat16_inputs_and_device_memory():
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray
    s, H = _ctx(24)
    nd, n = s.nd, 96
    rng = np.random.default_rng(1)
    A = bf16_round(rng.uniform(-1, 1, (nd, n)))
    B = bf16_round(rng.uniform(-1, 1, (nd, n)))
    ld = ((nd + 63) // 64) * 64
    Ap = np.zeros((n, ld), np.float32); Ap[:, :nd] = A.T
    Bp = np.zeros((n, ld), np.float32); Bp[:, :nd] = B.T
    dA, dB = DeviceArray.from_numpy(Ap), DeviceArray.from_numpy(Bp)
    L.check(L.lib().lynx_set_gram_precision(H.ctx, L.GRAM_BF16X3), H.ctx)
    c = np.zeros((n, n))
    L.check(L.lib().lynx_gram(H.ctx, L.F32, dA.value, ld, dB.value, ld, n, L.dptr(c), L.MEM_DEVICE), H.ctx)
    ref = s.latvec[0, 0] ** 3 / nd * (A.astype(np.float64).T @ B.astype(np.float64))
    assert np.abs(c.T - ref).max() / np.abs(ref).max() < 2e-6
    dA.free(); dB.free()


@pytest.mark.parametrize("grid,n", [(16, 24), ((20, 21, 22), 300), ((24, 25, 26), 129)])
def test_fp64_gram_matches_numpy(grid, n):
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray, synchronize
    s = System(grid, 10.0, np.zeros((0, 3)), with_nonlocal=False)
    H = s.hamiltonian(0)
    nd = s.nd
    rng = np.random.default_rng(2 + n)
    A = np.asfortranarray(rng.uniform(-0.5, 0.5, (nd, n)))
    B = np.asfortranarray(rng.uniform(-0.5, 0.5, (nd, n)))
    dV = abs(np.linalg.det(s.latvec)) / nd
    for X, Y in ((A, A), (A, B)):
        c = np.zeros((n, n))
        L.check(L.lib().lynx_gram(H.ctx, L.F64, L.dptr(X), nd, L.dptr(Y), nd, n, L.dptr(c), L.MEM_HOST), H.ctx)
        ref = dV * (X.T @ Y)
        assert np.abs(c.T - ref).max() / np.abs(ref).max() < 1e-13, (grid, n, X is Y)
    # device input, the overlap
    dA = DeviceArray.from_numpy(np.ascontiguousarray(A.T))
    c = np.zeros((n, n))
    L.check(L.lib().lynx_gram(H.ctx, L.F64, dA.value, nd, dA.value, nd, n, L.dptr(c), L.MEM_DEVICE), H.ctx)
    synchronize()
    dA.free()
    ref = dV * (A.T @ A)
    assert np.abs(c.T - ref).max() / np.abs(ref).max() < 1e-13


def test_gram_precision_rejects_unknown_mode():
    from lynx_amd import _lib as L
    _, H = _ctx(16)
    assert L.lib().lynx_set_gram_precision(H.ctx, 7) == L.E_ARG
