"""The SCF eigen-stage contract (lynx_scf_eig_*) against the oracle's
restatement of SCF::solve_eigenproblem and estimate_spectral_bounds:
rho_trigger passes in iteration 0, bound updates after every pass, Fermi
level and occupations; and the FP32-until-residual filter policy."""
import numpy as np
import pytest

import oracle
from systems import si8

pytestmark = pytest.mark.gpu

NB, NE, BETA = 20, 32.0, 1.0 / (3.166811563e-6 * 1000.0)   # Fermi-Dirac at 1000 K


@pytest.fixture(scope="module")
def si8_16():
    s = si8(16)
    return s


def test_eigen_stage_matches_oracle_fp64(si8_16):
    from lynx_amd import core
    s = si8_16
    op, H = s.oracle_op(), s.hamiltonian()
    X0 = s.random_block(NB, 81)
    ref = oracle.ScfEigStage(op, NB, NE, BETA)
    drv = core.ScfEigenStage(H, NB, NE, BETA)
    psi_ref = [X0.copy(order="F")]
    psi = [X0.copy(order="F")]
    for it in range(3):
        psi_ref, e_ref = ref.step(it, s.veff, psi_ref)
        e, occ, ef, prec = drv.step(it, 1.0, s.veff, psi)
        assert prec == 0
        assert np.abs(e - e_ref).max() < 1e-9 * max(1, np.abs(e_ref).max()), it
        lc, mn, mx = drv.state()
        assert abs(lc - ref.lambda_c) < 1e-9 and abs(mn[0] - ref.eig_min[0]) < 1e-9 and abs(mx[0] - ref.eig_max[0]) < 1e-9
        assert abs(ef - ref.fermi) < 1e-9
        assert np.abs(occ - ref.occ).max() < 1e-7


@pytest.mark.parametrize("gram", ["fp32", "bf16x3", "bf16x3_rot"])
def test_eigen_stage_fp32_until_residual(si8_16, gram):
    # the reduced-precision pass with FP32 Gram products and rotations, with
    # split-BF16x3 Gram products, and with split-BF16x3 rotations too
    from lynx_amd import core
    s = si8_16
    H64, H32 = s.hamiltonian(), s.hamiltonian()
    X0 = s.random_block(NB, 82)
    d64 = core.ScfEigenStage(H64, NB, NE, BETA)
    dmx = core.ScfEigenStage(H32, NB, NE, BETA, fp64_switch_residual=1e-4, gram=gram)
    p64, pmx = [X0.copy(order="F")], [X0.copy(order="F")]
    for it, res in enumerate((1.0, 1e-3, 1e-5, 1e-6)):
        e64, _, ef64, pr64 = d64.step(it, res, s.veff, p64)
        emx, _, efmx, prmx = dmx.step(it, res, s.veff, pmx)
        assert pr64 == 0 and prmx == (1 if res > 1e-4 or it == 0 else 0)
        # the FP32 filter perturbs the subspace at the 1e-6 level; the FP64
        # Rayleigh-Ritz keeps the eigenvalues within 1e-5 Ha of the FP64 run
        assert np.abs(emx - e64).max() < 1e-5, it
        assert abs(efmx - ef64) < 1e-5


def test_eigen_stage_fp32_complex_kpoints(si8_16):
    # the mixed-precision pass in complex arithmetic (complex64 filter and
    # Nd-sized products, complex128 Cholesky and eigensolve) on two k-points
    from systems import A_SI
    from lynx_amd import core
    s = si8_16
    kp = 2 * np.pi / A_SI * np.array([[0.0, 0.0, 0.0], [0.25, -0.125, 0.375]])
    w = np.array([0.5, 0.5])
    H64, H32 = s.hamiltonian(), s.hamiltonian()
    X0 = s.random_block(NB, 83, complex_=True)
    d64 = core.ScfEigenStage(H64, NB, NE, BETA, kpts=kp, kpt_weights=w)
    dmx = core.ScfEigenStage(H32, NB, NE, BETA, kpts=kp, kpt_weights=w, fp64_switch_residual=1e-4)
    p64 = [X0.copy(order="F"), X0.copy(order="F")]
    pmx = [X0.copy(order="F"), X0.copy(order="F")]
    for it, res in enumerate((1.0, 1e-3)):
        e64, _, ef64, _ = d64.step(it, res, s.veff, p64)
        emx, _, efmx, prmx = dmx.step(it, res, s.veff, pmx)
        assert prmx == 1
        assert np.abs(emx - e64).max() < 1e-5, it
        assert abs(efmx - ef64) < 1e-5


@pytest.mark.parametrize("nspin", [1, 2])
def test_eigen_stage_kpoints_on_device_blocks(si8_16, nspin):
    # FP64 passes over three k-points on device blocks run as one batched
    # solve per spin and pass (lynx_batch_solve, one batch per spin):
    # eigenvalues, bounds, Fermi level and occupations against the oracle's
    # per-k-point stage, and one counted solve per k-point, spin and pass
    import ctypes
    from systems import A_SI
    from lynx_amd import core, _lib as L
    from lynx_amd.hipmem import DeviceArray
    s = si8_16
    kp = 2 * np.pi / A_SI * np.array([[0.0, 0.0, 0.0], [0.25, -0.125, 0.375], [0.5, 0.25, 0.0]])
    w = np.array([0.25, 0.5, 0.25])
    op, H = s.oracle_op(), s.hamiltonian()
    X0 = s.random_block(NB, 89, complex_=True)
    veff = np.concatenate([s.veff, s.veff + 0.05][:nspin])
    ref = oracle.ScfEigStage(op, NB, NE, BETA, kpts=kp, kpt_weights=w, nspin=nspin)
    drv = core.ScfEigenStage(H, NB, NE, BETA, kpts=kp, kpt_weights=w, nspin=nspin)
    nb = 3 * nspin
    psi_ref = [X0.copy(order="F") for _ in range(nb)]
    dev = [DeviceArray.from_numpy(np.ascontiguousarray(X0.T)) for _ in range(nb)]   # column-major blocks
    cnt0, cnt = ctypes.c_long(), ctypes.c_long()
    L.check(L.lib().lynx_solve_count(H.ctx, ctypes.byref(cnt0)), H.ctx)
    for it in range(3):
        psi_ref, e_ref = ref.step(it, veff, psi_ref)
        e, occ, ef, prec = drv.step_device(it, 1.0, veff, [d.value for d in dev], s.nd)
        assert prec == 0
        assert np.abs(e - e_ref).max() < 1e-9 * max(1, np.abs(e_ref).max()), it
        lc, mn, mx = drv.state()
        assert abs(lc - ref.lambda_c) < 1e-9 and np.abs(mn - ref.eig_min).max() < 1e-9
        assert abs(ef - ref.fermi) < 1e-9 and np.abs(occ - ref.occ).max() < 1e-7
    L.check(L.lib().lynx_solve_count(H.ctx, ctypes.byref(cnt)), H.ctx)
    assert cnt.value - cnt0.value == ref.solves == nb * (4 + 1 + 1)
    for d in dev:
        d.free()


def test_eigen_stage_kpoints_on_device_blocks_mixed(si8_16):
    # reduced-precision passes over k-points on device blocks: the batched
    # filter in complex64 and the subspace steps in complex128, within
    # 1e-5 Ha of the FP64 stage (host blocks, per-k-point loop); FP64 after
    # the switch
    from systems import A_SI
    from lynx_amd import core
    from lynx_amd.hipmem import DeviceArray
    s = si8_16
    kp = 2 * np.pi / A_SI * np.array([[0.0, 0.0, 0.0], [0.25, -0.125, 0.375], [0.5, 0.25, 0.0]])
    w = np.array([0.25, 0.5, 0.25])
    H64, Hmx = s.hamiltonian(), s.hamiltonian()
    X0 = s.random_block(NB, 90, complex_=True)
    d64 = core.ScfEigenStage(H64, NB, NE, BETA, kpts=kp, kpt_weights=w)
    dmx = core.ScfEigenStage(Hmx, NB, NE, BETA, kpts=kp, kpt_weights=w, fp64_switch_residual=1e-4)
    p64 = [X0.copy(order="F") for _ in range(3)]
    dev = [DeviceArray.from_numpy(np.ascontiguousarray(X0.T)) for _ in range(3)]
    for it, res in enumerate((1.0, 1e-3, 1e-5)):
        e64, _, ef64, _ = d64.step(it, res, s.veff, p64)
        emx, _, efmx, prmx = dmx.step_device(it, res, s.veff, [d.value for d in dev], s.nd)
        assert prmx == (1 if res > 1e-4 or it == 0 else 0)
        assert np.abs(emx - e64).max() < 1e-5, it
        assert abs(efmx - ef64) < 1e-5
    for d in dev:
        d.free()


def test_eigen_stage_two_spins_matches_oracle(si8_16):
    # nspin = 2 with spin 1's Veff shifted by a constant: per-spin Lanczos
    # bounds, lambda_c from the larger last eigenvalue over spins, eig_min
    # updates after the first iteration, the pass count (rho_trigger, then
    # nchefsi, over both spins) through the context's solve counter
    import ctypes
    from lynx_amd import core, _lib as L
    s = si8_16
    op, H = s.oracle_op(), s.hamiltonian()
    veff = np.concatenate([s.veff, s.veff + 0.05])
    X0 = s.random_block(NB, 84)
    ref = oracle.ScfEigStage(op, NB, NE, BETA, nspin=2, rho_trigger=3, nchefsi=2)
    drv = core.ScfEigenStage(H, NB, NE, BETA, nspin=2, rho_trigger=3, nchefsi=2)
    psi_ref = [X0.copy(order="F"), X0.copy(order="F")]
    psi = [X0.copy(order="F"), X0.copy(order="F")]
    cnt = ctypes.c_long()
    for it in range(3):
        psi_ref, e_ref = ref.step(it, veff, psi_ref)
        e, occ, ef, prec = drv.step(it, 1.0, veff, psi)
        assert np.abs(e - e_ref).max() < 1e-9 * max(1, np.abs(e_ref).max()), it
        lc, mn, mx = drv.state()
        assert abs(lc - ref.lambda_c) < 1e-9
        assert np.abs(mn - ref.eig_min).max() < 1e-9 and np.abs(mx - ref.eig_max).max() < 1e-9
        assert abs(ef - ref.fermi) < 1e-9 and np.abs(occ - ref.occ).max() < 1e-7
        assert abs((ref.weights * occ).sum() - NE) < 1e-9
        L.check(L.lib().lynx_solve_count(H.ctx, ctypes.byref(cnt)), H.ctx)
        assert cnt.value == ref.solves == 6 + 4 * it


def test_wrapper_uploads_veff_again_after_the_stage(si8_16):
    # the eigen stage sets each spin's Veff on the context itself; a later
    # H.apply with the array the wrapper uploaded before must not run on the
    # stage's last (spin-1) potential, and the stage leaves the context's
    # filter precision as the caller set it
    import ctypes
    from lynx_amd import core, _lib as L
    s = si8_16
    op, H = s.oracle_op(), s.hamiltonian()
    vup = s.veff.copy()
    vdn = s.veff + 0.3
    X = s.random_block(3, 85)
    y0 = H.apply(H.halo, X.T.copy(), vup).reshape(3, -1).T          # caches vup in the wrapper
    L.check(L.lib().lynx_set_filter_precision(H.ctx, L.FILTER_BF16), H.ctx)
    drv = core.ScfEigenStage(H, NB, NE, BETA, nspin=2, rho_trigger=1, fp64_switch_residual=1e-4)
    psi = [s.random_block(NB, 86), s.random_block(NB, 87)]
    drv.step(0, 1.0, np.concatenate([vup, vdn]), psi)
    p = ctypes.c_int()
    L.check(L.lib().lynx_get_filter_precision(H.ctx, ctypes.byref(p)), H.ctx)
    assert p.value == L.FILTER_BF16
    y1 = H.apply(H.halo, X.T.copy(), vup).reshape(3, -1).T
    ref = op.apply(X, vup)
    assert rel(y1, ref) < 1e-12 and rel(y0, ref) < 1e-12


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
