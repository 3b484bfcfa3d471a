"""Batched CheFSI pass (lynx_batch_solve): the batch filter, batched
Cholesky-QR, one batched H psi step and batched Rayleigh-Ritz, against the
oracle's solve (the restatement of EigenSolver::solve / solve_kpt,
EigenSolver.cpp:738-842) of each unit alone, same start block and bounds:
Ritz values to 1e-9 relative, the returned block orthonormal
(dV X^H X = I to 1e-10) and spanning the oracle's subspace (singular values
of dV X^H X_oracle within 1e-8 of 1), lambda_cutoff = eig[N-1] + 0.1.
k-points of one FCC cell (complex128, one group), and a heterogeneous
real batch (three grids, band counts 10 / 6 / 10, two units in one group).
With the FP32 filter (lynx_batch_set_filter_precision): Ritz values within
1e-5 Ha of the oracle's FP64 solve, the block orthonormal to 1e-10 and its
subspace within 1e-4 of the oracle's; a precision other than FP64 or FP32
refused."""
import numpy as np
import pytest

from systems import System, A_SI, DIAMOND, displace, si2_fcc, si8

pytestmark = pytest.mark.gpu


def kvec(s, red):
    L = np.sqrt((s.latvec ** 2).sum(axis=1))
    return tuple(2 * np.pi * np.asarray(red) / L)


def _check(units, blocks, deg, bounds, dtype, filter_precision="fp64"):
    from lynx_amd import core
    from lynx_amd.hipmem import DeviceArray
    exact = filter_precision == "fp64"
    bf = core.BatchFilter([(H, k) for _, H, k in units])
    bf.set_filter_precision(filter_precision)
    dev = [DeviceArray.from_numpy(np.ascontiguousarray(b.T)) for b in blocks]
    lc, lo, hi = (np.array(v) for v in zip(*bounds))
    eigs, lco = bf.solve_device(dtype, [d.value for d in dev], [b.shape[0] for b in blocks],
                                [b.shape[1] for b in blocks], deg, lc, lo, hi)
    for i, (s, H, k) in enumerate(units):
        X = np.zeros((blocks[i].shape[1], blocks[i].shape[0]), dtype)
        dev[i].download(X)
        X = X.T
        op = s.oracle_op()
        if k is not None:
            op.set_kpoint(k)
        rpsi, reig, _ = op.solve(blocks[i], s.veff, lc[i], lo[i], hi[i], deg)
        dV = abs(np.linalg.det(s.latvec)) / s.nd
        assert np.abs(eigs[i] - reig).max() < (1e-9 * max(1.0, np.abs(reig).max()) if exact else 1e-5), i
        assert abs(lco[i] - (eigs[i][-1] + 0.1)) < 1e-14
        G = dV * X.conj().T @ X
        assert np.abs(G - np.eye(len(G))).max() < 1e-10, i
        sv = np.linalg.svd(dV * X.conj().T @ rpsi, compute_uv=False)
        assert np.abs(sv - 1).max() < (1e-8 if exact else 1e-4), (i, sv)
    for d in dev:
        d.free()
    bf.close()


def test_kpoint_batch_solve_matches_oracle():
    s = si2_fcc(16)
    H = s.hamiltonian()
    H.set_veff(s.veff)
    reds = [(0.0, 0.0, 0.0), (0.25, 0.0, 0.0), (0.25, 0.25, 0.0), (0.5, 0.25, -0.25)]
    ks = [kvec(s, r) for r in reds]
    units = [(s, H, k) for k in ks]
    blocks = [s.random_block(8, 70 + i, complex_=True) for i in range(len(ks))]
    bounds = [(0.3, -0.6, 60.0)] * len(ks)
    _check(units, blocks, 8, bounds, np.complex128)


def test_heterogeneous_real_batch_solve_matches_oracle():
    a = si8(16)
    b = System(16, A_SI, displace(DIAMOND, A_SI, 0.05, seed=9))
    c = si2_fcc(14)
    units = []
    for s in (a, b, c):
        H = s.hamiltonian()
        H.set_veff(s.veff)
        units.append((s, H, None))
    blocks = [a.random_block(10, 81), b.random_block(10, 82), c.random_block(6, 83)]
    bounds = [(0.3, -0.6, 60.0), (0.35, -0.62, 61.0), (0.2, -0.5, 70.0)]
    _check(units, blocks, 10, bounds, np.float64)


@pytest.mark.parametrize("kind", ["kpoints", "real"])
def test_batch_solve_with_fp32_filter(kind):
    if kind == "kpoints":
        s = si2_fcc(16)
        H = s.hamiltonian()
        H.set_veff(s.veff)
        ks = [kvec(s, r) for r in [(0.0, 0.0, 0.0), (0.25, 0.0, 0.0), (0.5, 0.25, -0.25)]]
        units = [(s, H, k) for k in ks]
        blocks = [s.random_block(8, 90 + i, complex_=True) for i in range(len(ks))]
        _check(units, blocks, 8, [(0.3, -0.6, 60.0)] * len(ks), np.complex128, "fp32")
    else:
        a, b = si8(16), si2_fcc(14)
        units = []
        for s in (a, b):
            H = s.hamiltonian()
            H.set_veff(s.veff)
            units.append((s, H, None))
        blocks = [a.random_block(10, 91), b.random_block(6, 92)]
        _check(units, blocks, 10, [(0.3, -0.6, 60.0), (0.2, -0.5, 70.0)], np.float64, "fp32")


def test_batch_solve_rejects_single_precision():
    from lynx_amd import _lib as L, core
    s = si8(12)
    H = s.hamiltonian()
    H.set_veff(s.veff)
    bf = core.BatchFilter([(H, None)])
    with pytest.raises(L.LynxError) as e:
        bf.solve_device(np.float32, [0], [s.nd], [4], 4, [0.3], [-0.6], [60.0])
    assert e.value.status == L.E_ARG
    assert L.lib().lynx_batch_set_filter_precision(bf._b, L.FILTER_BF16) == L.E_ARG
    bf.close()
