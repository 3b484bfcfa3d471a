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
        assert abs(lc - ref.lambda_c) < 1e-9 and abs(mn[0] - ref.eig_min) < 1e-9 and abs(mx[0] - ref.eig_max) < 1e-9
        assert abs(ef - ref.fermi) < 1e-9
        assert np.abs(occ - ref.occ).max() < 1e-7


def test_eigen_stage_fp32_until_residual(si8_16):
    from lynx_amd import core
    s = si8_16
    H64, H32 = s.hamiltonian(), s.hamiltonian()
    X0 = s.random_block(NB, 82)
    d64 = core.ScfEigenStage(H64, NB, NE, BETA)
    dmx = core.ScfEigenStage(H32, NB, NE, BETA, fp64_switch_residual=1e-4)
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
