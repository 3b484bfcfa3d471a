"""Atomic forces on the GPU (lynx_scf_forces) against the oracle's
restatement of Forces::compute (oracle/scf.py, pinned by finite
differences of the energy in tests/test_oracle_forces.py), after the same
SCF on both: the total (mean removed) and the local and nonlocal terms, to
1e-7 Ha/Bohr:

- displaced Si8 (C.pot stand-in), 16^3, PBE, Gamma point (real orbitals);
- BeCu with collinear spin, LDA PZ81, Gaussian smearing, two k-points
  (complex orbitals, Bloch phases in the gradient and the projectors);
- a displaced pair in the FCC primitive cell (non-orthogonal: lattice-
  direction gradients and the map to Cartesian components).
No psp8 with a model core charge ships with the reference, so the
core-correction term runs only in the oracle's pins.
"""
import os

import numpy as np
import pytest

import oracle
import oracle.scf as osc
from systems import A_SI, DIAMOND, FCC_PRIMITIVE, ROOT, psp

pytestmark = pytest.mark.gpu

KB = 3.1668115634556e-6
BECU = os.path.join(ROOT, "tests", "golden", "becu")


def _ham(lv, n, orth, files, types, frac):
    from lynx_amd import core
    lat = core.Lattice(lv, core.CellType.Orthogonal if orth else core.CellType.NonOrthogonal)
    grid = core.FDGrid(*n, lat)
    st = core.FDStencil(12, grid, lat)
    dom = core.Domain(grid)
    vnl = core.NonlocalProjector().setup_atoms(files, types, frac, True)
    return core.Hamiltonian(0).setup(st, dom, grid, core.HaloExchange(dom, 6), vnl)


def _case(name):
    if name == "si8_displaced_pbe":
        frac = DIAMOND + np.random.default_rng(7).normal(0, 0.08, DIAMOND.shape) / A_SI
        lv, n, orth, files, types = np.diag([A_SI] * 3), (16, 16, 16), True, [psp("C.pot")], [0] * 8
        return lv, n, orth, files, types, frac, dict(nband=24, beta=1 / (KB * 1000), xc=2), {}
    if name == "becu_spin_kpoints":
        a = 5.080050872061664
        frac = np.array([[0.04, 0.0, -0.05], [0.5, 0.53, 0.5]])
        ks = [(0.0, 0.0, 0.0), (0.0, 0.0, np.pi / a)]
        files = [os.path.join(BECU, "Be.pot"), os.path.join(BECU, "Cu.pot")]
        kw = dict(nband=18, beta=1 / (KB * 1160.45), xc=0, nspin=2, kpts=ks, kpt_weights=[0.5, 0.5], smearing=1)
        return np.diag([a] * 3), (21, 21, 21), True, files, [0, 1], frac, kw, {}
    lv = 7.0 * FCC_PRIMITIVE
    frac = np.array([[0.02, 0.01, 0.03], [0.27, 0.24, 0.22]])
    return lv, (16, 16, 16), False, [psp("C.pot")], [0, 0], frac, dict(nband=8, beta=1 / (KB * 1000), xc=0), {}


@pytest.mark.parametrize("name", ["si8_displaced_pbe", "becu_spin_kpoints", "fcc_pair"])
def test_forces_match_oracle(name):
    from lynx_amd.scf import SCF
    lv, n, orth, files, types, frac, kw, _ = _case(name)
    H = _ham(lv, n, orth, files, types, frac)
    nband, beta = kw.pop("nband"), kw.pop("beta")
    xc = kw.pop("xc")
    S = SCF(H, nband, beta, xc=xc, tol=1e-8, solver="spectral", **kw)
    got = S.run()
    F, parts = S.forces(with_parts=True)
    og = oracle.Grid(*n, oracle.Lattice(lv, orth))
    O = osc.ScfOracle(og, [oracle.load_psp8(f) for f in files], types, frac, nband, beta, xc=xc, tol=1e-8,
                      solver="spectral", **kw)
    ref = O.run()
    Fo, po = O.forces()
    assert got["converged"] and ref["converged"] and got["iterations"] == len(O.history)
    print(f"[forces] {name}: max |F| {np.abs(Fo).max():.4f} Ha/Bohr, GPU - oracle {np.abs(F - Fo).max():.2e}",
          flush=True)
    assert np.abs(Fo).max() > 5e-3
    assert np.abs(F - Fo).max() < 1e-7
    for a, b in zip(parts[:2], po[:2]):
        assert np.abs(a - b).max() < 1e-7
    assert np.abs(F.sum(axis=0)).max() < 1e-12
    S.close()


def test_forces_before_a_run_is_a_state_error():
    from lynx_amd import _lib as L
    from lynx_amd.scf import SCF
    lv, n, orth, files, types, frac, kw, _ = _case("fcc_pair")
    S = SCF(_ham(lv, n, orth, files, types, frac), 8, 1 / (KB * 1000), xc=0)
    with pytest.raises(L.LynxError) as e:
        S.forces()
    assert e.value.status == L.E_STATE
    S.close()
