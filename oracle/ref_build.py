"""Compiles the buildable subset of the reference (read in place under
/root/reference, never copied) with oracle/ref_driver.cpp into
oracle/_ref/libref.so.  Test infrastructure only; a missing reference tree is
not an error (machines that only run the built library have none)."""
import os
import subprocess
import sys

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
SOURCES = ["core/Lattice.cpp", "core/FDGrid.cpp", "core/Domain.cpp", "core/KPoints.cpp",
           "core/NumericalMethods.cpp", "operators/FDStencil.cpp", "parallel/HaloExchange.cpp",
           "atoms/Pseudopotential.cpp", "electronic/Wavefunction.cpp", "electronic/Occupation.cpp",
           "parallel/MPIComm.cpp", "atoms/AtomType.cpp", "atoms/Crystal.cpp", "core/ParameterDefaults.cpp",
           "physics/Electrostatics.cpp"]
# MPICH (Occupation's communicator type; no MPI call is made with null
# communicators) and the MKL runtime (LAPACK dsterf_ for the Lanczos
# template) from the container's conda tree
CONDA = "/opt/conda"


def main():
    if not os.path.isdir(REF):
        print("reference tree absent; oracle/_ref not built", file=sys.stderr)
        return 0
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(REF, s) for s in SOURCES] + [os.path.join(HERE, "ref_driver.cpp")]
    cmd = (["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", REF, "-I", os.path.join(CONDA, "include"),
            "-o", os.path.join(OUT, "libref.so")] + srcs +
           ["-L", os.path.join(CONDA, "lib"), "-lmkl_rt", "-lmpi",
            # the system's libstdc++ (the one g++ compiled against) is found
            # before the conda tree's older copy
            "-Wl,-rpath,/usr/lib/x86_64-linux-gnu", "-Wl,-rpath," + os.path.join(CONDA, "lib")])
    return subprocess.run(cmd).returncode


if __name__ == "__main__":
    sys.exit(main())
