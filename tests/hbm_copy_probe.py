"""Calibration run for the profiler's HBM byte counters (not a test module):
a device-to-device copy of NBYTES, repeated 5 times, whose FETCH_SIZE and
WRITE_SIZE readings give the scale factors applied to the stencil kernels'
counter summaries.

    python tests/hbm_copy_probe.py 268435456
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from lynx_amd.hipmem import D2D, DeviceArray, _ok, hip, synchronize  # noqa: E402


def main(nbytes):
    a = DeviceArray(nbytes)
    b = DeviceArray(nbytes)
    a.fill_bytes(1)
    for _ in range(5):
        _ok(hip().hipMemcpy(b.ptr, a.ptr, nbytes, D2D), "hipMemcpy D2D")
    synchronize()
    print(f"copied {nbytes} bytes x5")


if __name__ == "__main__":
    if len(sys.argv) != 2:
        sys.exit("usage: python tests/hbm_copy_probe.py NBYTES")
    main(int(sys.argv[1]))
