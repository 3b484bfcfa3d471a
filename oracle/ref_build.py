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
           "atoms/Pseudopotential.cpp"]


def main():
    if not os.path.isdir(REF):
        print("reference tree absent; oracle/_ref not built", file=sys.stderr)
        return 0
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(REF, s) for s in SOURCES] + [os.path.join(HERE, "ref_driver.cpp")]
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", REF, "-o", os.path.join(OUT, "libref.so")] + srcs
    return subprocess.run(cmd).returncode


if __name__ == "__main__":
    sys.exit(main())
