"""The reference's density restart file and result JSON (lynx_amd/io.py),
on the CPU:

- tests/golden/density_batio3 holds a density file the reference wrote for
  its BaTiO3 job (15^3 grid, one spin) with that job's input.json.  Its header
  matches the job's grid and lattice, dV is the cell volume over Nx Ny Nz,
  the density integrates to the job's 40 valence electrons (Ba 10, Ti 12,
  3 x O 6, the zion in its psp8 file names), and writing the same density,
  grid, lattice and dV gives the file byte for byte (the file carries the
  earlier "SPARCRHO" tag; the current reader takes "LYNX_RHO" only, so the
  default read refuses it as the reference's DensityIO::read would);
- every refusal of DensityIO::read (DensityIO.cpp:51-121): missing or short
  file, tag, version, grid, spin count;
- the result JSON against tests/unit/test_ResultWriter.cpp's two cases.
"""
import json
import os
import struct

import numpy as np
import pytest

from lynx_amd import io
from systems import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "density_batio3")
BIN = os.path.join(GOLD, "electron_density.bin")


@pytest.fixture(scope="module")
def job():
    with open(os.path.join(GOLD, "input.json")) as f:
        return json.load(f)


def test_reference_density_file_header(job):
    h = io.density_header(BIN)
    g = job["grid"]
    assert h["magic"] == b"SPARCRHO" and h["version"] == 1 and h["nspin"] == 1
    assert h["shape"] == (g["Nx"], g["Ny"], g["Nz"])
    lv = np.array(job["lattice"]["vectors"], float)
    assert np.array_equal(h["latvec"], lv)
    assert abs(h["dV"] - abs(np.linalg.det(lv)) / np.prod(h["shape"])) < 1e-15 * h["dV"]
    assert os.path.getsize(BIN) == 128 + 8 * int(np.prod(h["shape"]))


def test_reference_density_file_reads_and_rewrites_bytewise(job):
    g = job["grid"]
    shape = (g["Nx"], g["Ny"], g["Nz"])
    assert io.read_density(BIN, shape, 1) is None          # the current tag only
    rho = io.read_density(BIN, shape, 1, magic=b"SPARCRHO")
    assert rho.shape == (1, 15 ** 3)
    h = io.density_header(BIN)
    # Ne = 10 (Ba) + 12 (Ti) + 3 x 6 (O)
    assert abs(rho.sum() * h["dV"] - 40.0) < 1e-12
    ne = sum(int(os.path.basename(a["pseudo_file"]).split("_")[2]) * len(a["coordinates"]) for a in job["atoms"])
    assert ne == 40
    out = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"lynx_rho_{os.getpid()}.bin")
    try:
        assert io.write_density(out, rho, shape, h["latvec"], h["dV"])
        with open(out, "rb") as f:
            mine = f.read()
        with open(BIN, "rb") as f:
            ref = f.read()
        assert mine[:8] == b"LYNX_RHO" and mine[8:] == ref[8:]
        assert np.array_equal(io.read_density(out, shape, 1), rho)
    finally:
        if os.path.exists(out):
            os.remove(out)


def test_two_spin_round_trip_and_totals(tmp_path):
    rng = np.random.default_rng(2)
    shape = (4, 5, 6)
    up, dn = rng.uniform(0, 1, 120), rng.uniform(0, 1, 120)
    p = str(tmp_path / "rho2.bin")
    lv = np.array([[4.0, 0.1, 0.0], [0.0, 5.0, 0.2], [0.3, 0.0, 6.0]])
    assert io.write_density(p, np.stack([up, dn]), shape, lv, 0.5)
    got = io.read_density(p, shape, 2)
    assert np.array_equal(got[0], up) and np.array_equal(got[1], dn)
    t, m = io.total_and_mag(got)
    assert np.array_equal(t, up + dn) and np.array_equal(m, up - dn)
    t1, m1 = io.total_and_mag(got[:1])
    assert np.array_equal(t1, up) and m1 is None
    h = io.density_header(p)
    # header offsets: nspin at 12, grid at 16..27, latvec row-major at 32, dV at 104
    raw = open(p, "rb").read(128)
    assert struct.unpack_from("<I", raw, 12)[0] == 2
    assert struct.unpack_from("<3I", raw, 16) == shape
    assert struct.unpack_from("<9d", raw, 32) == tuple(lv.ravel())
    assert struct.unpack_from("<d", raw, 104)[0] == 0.5 and raw[112:] == bytes(16)
    assert h["nspin"] == 2


def test_reader_refusals(tmp_path):
    shape = (3, 3, 3)
    p = str(tmp_path / "r.bin")
    assert io.write_density(p, np.ones(27), shape, np.eye(3), 1.0)
    assert io.read_density(p, shape, 1) is not None
    assert io.read_density(str(tmp_path / "missing.bin"), shape, 1) is None
    assert io.read_density(p, (3, 3, 4), 1) is None        # grid
    assert io.read_density(p, (3, 3, 3), 2) is None        # spin count
    raw = bytearray(open(p, "rb").read())
    bad = str(tmp_path / "bad.bin")
    for off, val in ((0, b"LYNX_RHX"), (8, struct.pack("<I", 2))):   # tag, version
        b = bytearray(raw)
        b[off:off + len(val)] = val
        open(bad, "wb").write(bytes(b))
        assert io.read_density(bad, shape, 1) is None
    open(bad, "wb").write(bytes(raw[:-8]))                   # short body
    assert io.read_density(bad, shape, 1) is None
    open(bad, "wb").write(bytes(raw[:100]))                  # short header
    assert io.read_density(bad, shape, 1) is None
    with pytest.raises(ValueError):
        io.write_density(p, np.ones(26), shape, np.eye(3), 1.0)


def test_result_writer_minimal(tmp_path):
    # tests/unit/test_ResultWriter.cpp:12-34
    rw = io.ResultWriter()
    rw.test_name, rw.input_file, rw.device, rw.mpi_ranks, rw.lynx_version = "Unit.Minimal", "dummy.json", "CPU", 1, "test"
    p = str(tmp_path / "sub" / "min.json")
    rw.write(p)
    j = json.load(open(p))
    assert j["test_name"] == "Unit.Minimal" and j["input_file"] == "dummy.json" and j["device"] == "CPU"
    assert j["mpi_ranks"] == 1 and j["lynx_version"] == "test" and "timestamp_utc" in j
    # every section present with the reference's defaults
    assert j["grid"]["boundary_conditions"] == ["P", "P", "P"] and j["electronic"]["spin"] == "none"
    assert j["kpoints"]["grid"] == [1, 1, 1] and j["kpoints"]["n_irreducible"] == 0
    assert j["stress"]["computed"] is False and j["eigenvalues"]["units"] == "Ha"
    assert set(j) == {"lynx_version", "test_name", "input_file", "timestamp_utc", "device", "mpi_ranks", "system",
                      "cell", "grid", "electronic", "kpoints", "scf", "energies_Ha", "forces_Ha_per_bohr", "stress",
                      "eigenvalues", "timing_s"}
    text = open(p).read()
    assert text.startswith('{\n  "cell"')      # two-space indent, sorted keys


def test_result_writer_all_sections(tmp_path):
    # tests/unit/test_ResultWriter.cpp:36-170, the same values and checks
    rw = io.ResultWriter()
    rw.test_name, rw.input_file, rw.lynx_version = "Unit.AllSections", "in.json", "test"
    rw.set_system(n_atoms=2, n_electrons=8, atoms_cartesian_bohr=[[0, 0, 0], [1, 1, 1]],
                  atoms_fractional=[[0, 0, 0], [0.5, 0.5, 0.5]],
                  atom_types=[{"element": "Si", "Zval": 4.0, "mass_amu": 28.0855, "n_atoms": 2,
                               "pseudo_file": "/path/Si.psp8"}])
    rw.set_cell(lattice_vectors_bohr=[[2, 0, 0], [0, 2, 0], [0, 0, 2]], volume_bohr3=8.0, orthogonal=True)
    rw.set_grid(Nx=10, Ny=10, Nz=10, dx_bohr=0.2, dy_bohr=0.2, dz_bohr=0.2, fd_order=12)
    rw.set_electronic(xc="PBE", spin="none", soc=False, smearing="gaussian", electronic_temperature_K=315.775,
                      kBT_Ha=1e-3, n_states=6, cheb_degree=20)
    rw.set_kpoints(grid=[1, 1, 1], shift=[0, 0, 0], list=[{"index": 1, "reduced": [0, 0, 0], "weight": 1.0}])
    rw.set_scf(max_iter=100, tolerance=1e-6, mixing_variable="density", preconditioner="kerker", mixing_history=7,
               mixing_parameter=0.3, converged=True, n_iterations=3,
               history=[{"iter": i + 1, "free_energy_per_atom_Ha": -3.87, "scf_error": e, "time_s": t}
                        for i, (e, t) in enumerate(((1.8e-1, 0.5), (2.3e-2, 0.4), (5.1e-7, 0.4)))])
    rw.set_energies(total_free_energy=-15.48, free_energy_per_atom=-7.74, band_structure=-1.89,
                    exchange_correlation=-4.43, hartree=0.0, self_energy=-50.0, correction_energy=-7.6,
                    self_plus_correction=-57.6, entropy_term=-0.017, fermi_level=0.008)
    rw.set_forces(max=0.023, rms=0.020, per_atom=np.array([[0.01, 0.02, 0.03], [-0.01, -0.02, -0.03]]))
    rw.set_stress(computed=True, voigt_Ha_per_bohr3=[1, 2, 3, 4, 5, 6], voigt_GPa=[1, 2, 3, 4, 5, 6],
                  tensor_Ha_per_bohr3=[[1, 2, 3], [2, 4, 5], [3, 5, 6]], tensor_GPa=[[1, 2, 3], [2, 4, 5], [3, 5, 6]],
                  pressure_GPa=1.0, max_abs_stress_GPa=6.0)
    rw.set_eigenvalues(n_states=2, per_kpoint=[{"kpoint_index": 1, "reduced": [0, 0, 0], "weight": 1.0,
                                                "spin_up": {"eigenvalues": [-0.5, -0.4], "occupations": [2.0, 2.0]}}])
    rw.set_timing(scf_total=5.0, force=0.1, stress=0.3, walltime_total=6.0)
    with pytest.raises(KeyError):
        rw.set_timing(total=1.0)
    p = str(tmp_path / "all.json")
    rw.write(p)
    j = json.load(open(p))
    assert j["system"]["n_atoms"] == 2 and j["cell"]["volume_bohr3"] == 8.0 and j["grid"]["fd_order"] == 12
    assert j["electronic"]["xc"] == "PBE" and len(j["kpoints"]["list"]) == 1 and len(j["scf"]["history"]) == 3
    assert j["energies_Ha"]["total_free_energy"] == -15.48 and len(j["forces_Ha_per_bohr"]["per_atom"]) == 2
    assert j["stress"]["computed"] is True and len(j["stress"]["voigt_GPa"]) == 6
    assert len(j["eigenvalues"]["per_kpoint"][0]["spin_up"]["eigenvalues"]) == 2
    assert "spin_down" not in j["eigenvalues"]["per_kpoint"][0]
    assert j["timing_s"]["walltime_total"] == 6.0 and j["kpoints"]["n_irreducible"] == 1
    # non-finite numbers become null, as the reference's JSON library writes them
    rw.set_energies(fermi_level=float("nan"))
    rw.write(p)
    assert json.load(open(p))["energies_Ha"]["fermi_level"] is None
