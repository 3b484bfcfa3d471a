"""The C++ drop-in header on the GPU: a compiled program using
lynx_mi355x::Hamiltonian / EigenSolver / ScfEigenStage gives the same
eigenvalues, bounds and Fermi level as the same calls through the Python API
(both run the library's FP64 path, so they agree to rounding)."""
import os
import subprocess

import numpy as np
import pytest

from systems import ROOT, psp, si8

pytestmark = pytest.mark.gpu


def test_cpp_classes_match_python_api(tmp_path):
    from lynx_amd import core
    exe = tmp_path / "gpu_api"
    lib_dir = os.path.join(ROOT, "lynx_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "gpu_api.cpp"), "-L", lib_dir, "-llynx_chefsi",
                        f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    s = si8(16)
    X = s.random_block(20, 91)
    s.veff.astype(np.float64).tofile(tmp_path / "veff.bin")
    np.asfortranarray(X).ravel(order="F").tofile(tmp_path / "psi.bin")
    out = subprocess.run([str(exe), psp("C.pot"), str(tmp_path / "veff.bin"), str(tmp_path / "psi.bin")],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = {ln.split()[0]: np.array(ln.split()[1:], float) for ln in out.stdout.splitlines()}
    H = s.hamiltonian()
    es = core.EigenSolver().setup(s.context(20), H)
    lo, hi = es.lanczos_bounds(s.veff)
    assert np.allclose(lines["bounds"], [lo, hi], rtol=1e-12, atol=1e-12)
    _, eig = es.solve(X, s.veff, 0.5 * (lo + hi), lo, hi, cheb_degree=10)
    assert np.abs(lines["solve"] - eig).max() < 1e-10
    H2 = s.hamiltonian()
    stage = core.ScfEigenStage(H2, 20, 32.0, 315.77)
    e2, _, ef, _ = stage.step(0, 1.0, s.veff, [X.copy(order="F")])
    assert abs(lines["stage"][0] - ef) < 1e-10
    assert np.abs(lines["stage"][1:] - e2).max() < 1e-10
    # a NaN Veff is refused with an exception; the earlier array then works again
    assert lines["refused"][0] == 1
    assert np.allclose(lines["refused"][1:], [lo, hi], rtol=1e-12, atol=1e-12)
    # a change made in place to one Veff entry reaches the device
    err, ymax, step = lines["inplace"]
    assert step > 1e-3
    assert err < 1e-12 * max(1.0, ymax)
