"""Density restart and result files on the GPU, against the oracle:

- a converged density written to the reference's density file, read back
  bit for bit, and handed to a new SCF (lynx_scf_set_initial_density)
  starts the run the oracle starts from the same density
  (ScfOracle.run(rho0, mag0), SCF::set_initial_density): per-iteration
  energy and residual as in the whole-run parity tests, and the converged
  energy of the atomic start to 1e-7 Ha/atom (the orbitals still start
  random, so the iteration count need not drop);
  one spin (Si8 16^3 PBE, Gamma) and two spins with their magnetization
  (BeCu, two k-points, PZ81, Gaussian smearing);
- the Calculator's "output" keys: density_output_file written after the
  run, density_restart_file read before the next one (a file of another
  grid falls back to the atomic start, as SCF.cpp:742-761 does), and the
  result JSON holding the run's energies, history, eigenvalues and forces.
"""
import json
import os

import numpy as np
import pytest

import oracle.scf as osc
from lynx_amd import io
from systems import A_SI, DIAMOND, psp
from test_gpu_scf import KB, becu, si8

pytestmark = pytest.mark.gpu


def _compare(S, O, natom, rho0, mag0=None):
    S.set_initial_density(rho0, mag0)
    got = S.run()
    ref = O.run(rho0=rho0, mag0=mag0)
    h = S.history()
    assert got["converged"] and ref["converged"]
    assert len(h["etotal"]) == len(O.history), (len(h["etotal"]), len(O.history))
    for i, rec in enumerate(O.history):
        assert abs(h["etotal"][i] - rec["etotal"]) < 1e-7, (i, h["etotal"][i], rec["etotal"])
        assert abs(h["error"][i] - rec["error"]) < 1e-6 * rec["error"] + 1e-12, (i, h["error"][i], rec["error"])
    assert abs(got["etotal"] - ref["etotal"]) / natom < 1e-8
    return got


def _round_trip(tmp_path, rho, shape, nspin):
    p = str(tmp_path / "rho.bin")
    assert io.write_density(p, rho.reshape(nspin, -1), shape, np.eye(3), 1.0)
    back = io.read_density(p, shape, nspin)
    assert np.array_equal(back, rho.reshape(nspin, -1))
    return io.total_and_mag(back)


def test_restart_si8_one_spin_matches_oracle(tmp_path):
    from lynx_amd.scf import SCF
    H, og, ps, types, frac = si8(16)
    kw = dict(xc="GGA_PBE", tol=1e-6, solver="spectral")
    S0 = SCF(H, 24, 1 / (KB * 1000), **kw)
    first = S0.run()
    _, _, rho, _ = S0.state()
    S0.close()
    total, mag = _round_trip(tmp_path, rho, (16, 16, 16), 1)
    assert mag is None
    S = SCF(H, 24, 1 / (KB * 1000), **kw)
    O = osc.ScfOracle(og, ps, types, frac, 24, 1 / (KB * 1000), xc=2, tol=1e-6, solver="spectral")
    got = _compare(S, O, 8, total)
    print(f"[restart] Si8 16^3: {first['iterations']} iterations from atoms, {got['iterations']} from the "
          f"restart density; E {first['etotal']:.10f} / {got['etotal']:.10f} Ha", flush=True)
    assert abs(got["etotal"] - first["etotal"]) / 8 < 1e-7
    # rho None restores the atomic start: the first run again
    S.set_initial_density(None)
    again = S.run()
    assert again["iterations"] == first["iterations"] and abs(again["etotal"] - first["etotal"]) < 1e-9
    S.close()


def test_restart_becu_two_spins_matches_oracle(tmp_path):
    from lynx_amd.scf import SCF
    H, og, ps, types, frac, ks = becu(2)
    beta = 1 / (KB * 1160.45)
    kw = dict(xc="LDA_PZ", nspin=2, kpts=ks, kpt_weights=[0.5, 0.5], smearing=1, tol=1.58e-4, solver="spectral")
    S0 = SCF(H, 18, beta, **kw)
    first = S0.run()
    _, _, rho, _ = S0.state()
    S0.close()
    total, mag = _round_trip(tmp_path, rho, (21, 21, 21), 2)
    S = SCF(H, 18, beta, **kw)
    O = osc.ScfOracle(og, ps, types, frac, 18, beta, xc=0, nspin=2, kpts=ks, kpt_weights=[0.5, 0.5], smearing=1,
                      tol=1.58e-4, solver="spectral")
    got = _compare(S, O, 2, total, mag)
    print(f"[restart] BeCu spin: {first['iterations']} iterations from atoms, {got['iterations']} from the "
          f"restart density", flush=True)
    assert abs(got["etotal"] - first["etotal"]) / 2 < 1e-6
    S.close()


def _si8_config(n, out):
    return {"lattice": {"vectors": (np.eye(3) * A_SI).tolist(), "cell_type": "orthogonal"},
            "grid": {"Nx": n, "Ny": n, "Nz": n, "fd_order": 12},
            "atoms": [{"element": "Si", "pseudo_file": psp("C.pot"), "fractional": True,
                       "coordinates": DIAMOND.tolist()}],
            "electronic": {"xc": "GGA_PBE", "spin": "none", "temperature": 1000.0, "smearing": "fermi-dirac",
                           "Nstates": 24},
            "scf": {"tolerance": 1e-6, "cheb_degree": 20},
            "output": out}


def test_calculator_density_files_and_results(tmp_path):
    from lynx_amd.scf import Calculator
    rho_file = str(tmp_path / "si8_rho.bin")
    a = Calculator(0).set_config(_si8_config(14, {"density_output_file": rho_file}))
    ea = a.run()
    fa = a.compute_forces()
    h = io.density_header(rho_file)
    assert h["magic"] == io.DENSITY_MAGIC and h["shape"] == (14, 14, 14) and h["nspin"] == 1
    assert np.array_equal(h["latvec"], np.eye(3) * A_SI) and h["dV"] == a.dV
    rho = io.read_density(rho_file, (14, 14, 14), 1)
    assert np.array_equal(rho[0], a.density())
    assert abs(rho.sum() * a.dV - 32.0) < 1e-9

    b = Calculator(0).set_config(_si8_config(14, {"density_restart_file": rho_file}))
    eb = b.run()
    print(f"[restart] Calculator Si8 14^3: {a.n_iterations()} iterations, restart {b.n_iterations()}; "
          f"E {ea:.10f} / {eb:.10f}", flush=True)
    assert b.is_converged() and abs(eb - ea) / 8 < 1e-7
    assert abs(b.density().sum() * b.dV - 32.0) < 1e-9

    # a density file of another grid is refused: the atomic start, the first run again
    other = str(tmp_path / "other.bin")
    assert io.write_density(other, np.ones(12 ** 3), (12, 12, 12), np.eye(3) * A_SI, 1.0)
    c = Calculator(0).set_config(_si8_config(14, {"density_restart_file": other}))
    ec = c.run()
    assert c.n_iterations() == a.n_iterations() and abs(ec - ea) < 1e-9

    res = str(tmp_path / "out" / "lynx_results.json")
    a.write_results(res, test_name="Si8.Restart", input_file="si8.json")
    j = json.load(open(res))
    r = a.energy()
    assert j["test_name"] == "Si8.Restart" and j["device"] == "GPU"
    assert j["system"]["n_atoms"] == 8 and j["system"]["n_electrons"] == 32
    assert j["system"]["atom_types"][0]["Zval"] == 4.0 and j["system"]["atom_types"][0]["n_atoms"] == 8
    assert np.allclose(j["system"]["atoms_fractional"], DIAMOND, atol=1e-15)
    assert j["grid"]["Nx"] == 14 and abs(j["grid"]["dx_bohr"] - A_SI / 14) < 1e-15
    assert abs(j["cell"]["volume_bohr3"] - A_SI ** 3) < 1e-9 and j["cell"]["orthogonal"] is True
    assert j["electronic"]["xc"] == "PBE" and j["electronic"]["smearing"] == "fermi-dirac"
    assert abs(j["electronic"]["kBT_Ha"] - KB * 1000.0) < 1e-18
    assert j["kpoints"]["n_irreducible"] == 1 and j["kpoints"]["list"][0]["weight"] == 1.0
    e = j["energies_Ha"]
    assert e["total_free_energy"] == r["etotal"] and e["free_energy_per_atom"] == r["etotal"] / 8
    assert e["self_plus_correction"] == r["eself"] + r["ec"] and e["fermi_level"] == r["fermi"]
    s = j["scf"]
    assert s["converged"] is True and s["n_iterations"] == a.n_iterations() == len(s["history"])
    assert s["history"][-1]["free_energy_per_atom_Ha"] == r["etotal"] / 8
    assert s["history"][-1]["scf_error"] == r["scf_error"]
    f = j["forces_Ha_per_bohr"]
    assert np.array_equal(np.array(f["per_atom"]), fa)
    assert f["max"] == np.abs(fa).max() and abs(f["rms"] - np.sqrt((fa ** 2).mean())) < 1e-15
    ev = j["eigenvalues"]
    assert ev["n_states"] == 24 and len(ev["per_kpoint"]) == 1 and "spin_down" not in ev["per_kpoint"][0]
    eig, occ, _, _ = a.scf.state()
    assert ev["per_kpoint"][0]["spin_up"]["eigenvalues"] == eig.tolist()
    assert abs(sum(ev["per_kpoint"][0]["spin_up"]["occupations"]) * 2 - 32.0) < 1e-8
    assert j["timing_s"]["scf_total"] > 0 and j["stress"]["computed"] is False
    for x in (a, b, c):
        x.scf.close()
