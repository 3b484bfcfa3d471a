#!/usr/bin/env python
"""Chebyshev filter throughput on one MI355X.

Workload (filter-only operator bench): Si64, the 2x2x2 supercell of the
8-atom diamond cell (a = 20.52 Bohr) with atoms displaced by Gaussian
sigma = 0.05 Bohr (seed 1); 60^3 real grid, 12th-order FD stencil, nonlocal
projectors from the C.pot psp8 (stand-in for the absent Si psp8: zion 4, two
s and two p projectors); a block X of 160 real orbitals, uniform [-0.5, 0.5]
(seed 2); degree-20 Chebyshev filter over bounds from 10 Lanczos steps.
Veff is a Gaussian well per atom (stand-in for the atomic-superposition
potential; the filter's cost does not depend on its values).

One step = one degree-20 filter call over the whole block, inputs resident in
HBM.  value = B_filter / t_step with B_filter = s*P*(3m-1) + 8*Nd*m, the
algorithmic bytes of the fused recurrence (P = Nd*ncol point-columns,
s = bytes per element, m = degree).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "SCF wall-time & filter GB/s: FP32/BF16 batched vs FP64 at ≤1e-5 Ha/atom"
HBM_PEAK_GBS = 8000.0
N, A, NCOL, DEGREE, LANCZOS_STEPS = 60, 20.52, 160, 20, 10


def filter_bytes(nd, ncol, degree, esize):
    return esize * nd * ncol * (3 * degree - 1) + 8 * nd * degree


def profiled_traffic(shape, kernel_prefix):
    """HBM bytes per launch of the stencil kernel from a committed rocprofv3
    summary (profiles/*.md: FETCH_SIZE and WRITE_SIZE runs, calibrated on a
    device copy) for the same launch shape, or None."""
    import glob
    key = "%d,%d,%d,%d" % (shape["vector_runs"], shape["threads"], shape["points_div"], shape["z_chunks"])
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        fields = {}
        with open(path) as f:
            for line in f:
                if ":" in line and line.split(":", 1)[0] in ("traffic_bytes_per_launch", "kernel", "shape"):
                    k, v = line.split(":", 1)
                    fields[k] = v.strip()
        if fields.get("shape") == key and fields.get("kernel", "").startswith(kernel_prefix):
            return int(fields["traffic_bytes_per_launch"]), os.path.relpath(path, ROOT)
    return None, None


def build_system():
    from systems import System, displace, supercell
    frac = displace(supercell(2), A, 0.05, seed=1)
    return System(N, A, frac)


def run_gpu(sysm, dtype_code, steps, warmup, X64, shape=None):
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray, synchronize
    lib = L.lib()
    H = sysm.hamiltonian(0)
    H.set_veff(sysm.veff)
    ctx = H.ctx
    if shape is not None:
        vec, thr_, div_, zc_ = shape
        L.check(lib.lynx_set_stencil_config(ctx, thr_, div_, zc_, vec), ctx)
    lo = ctypes.c_double(); hi = ctypes.c_double(); it = ctypes.c_int()
    L.check(lib.lynx_lanczos_bounds(ctx, L.F64, ctypes.byref(lo), ctypes.byref(hi), 0.0, LANCZOS_STEPS, 1,
                                    ctypes.byref(it)), ctx)
    lam_c = 0.5 * (lo.value + hi.value)
    nd = sysm.nd
    Xh = np.asfortranarray(X64.astype(L.np_dtype(dtype_code)))
    Xd = DeviceArray.from_numpy(Xh.T.copy())
    Yd = DeviceArray(Xh.nbytes)
    args = (ctx, dtype_code, Xd.value, nd, Yd.value, nd, NCOL, DEGREE, lam_c, lo.value, hi.value, L.MEM_DEVICE)
    for _ in range(warmup):
        L.check(lib.lynx_cheb_filter(*args), ctx)
    L.check(lib.lynx_set_timing(ctx, 1), ctx)
    kern_ms = np.zeros(4)
    kern_n = np.zeros(4, dtype=int)
    t_ms = np.zeros(4)
    t_n = (ctypes.c_int * 4)()
    t_d = (ctypes.c_double * 4)()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        L.check(lib.lynx_cheb_filter(*args), ctx)
        L.check(lib.lynx_last_filter_timing(ctx, t_d, t_n), ctx)
        t_ms[:] = list(t_d)
        kern_ms += t_ms
        kern_n += np.array(list(t_n))
    synchronize()
    elapsed = time.perf_counter() - t0
    L.check(lib.lynx_set_timing(ctx, 0), ctx)
    thr = ctypes.c_int(); pd = ctypes.c_int(); zc = ctypes.c_int(); vr = ctypes.c_int(); nt = ctypes.c_int(64)
    tms = (ctypes.c_double * 64)()
    L.check(lib.lynx_get_stencil_config(ctx, dtype_code, NCOL, ctypes.byref(thr), ctypes.byref(pd),
                                        ctypes.byref(zc), ctypes.byref(vr), tms, ctypes.byref(nt)), ctx)
    Y = np.zeros((NCOL, nd), dtype=Xh.dtype)
    Yd.download(Y)
    finite = bool(np.all(np.isfinite(Y)))
    esize = Xh.dtype.itemsize
    B = filter_bytes(nd, NCOL, DEGREE, esize)
    ms_step = 1e3 * elapsed / steps
    stencil_ms = kern_ms[1] / steps
    stencil_launches = kern_n[1] / steps
    res = dict(
        gbs=B / (elapsed / steps) / 1e9,
        ms_per_step=ms_step,
        bytes_per_step=B,
        stencil_ms_per_step=stencil_ms,
        stencil_avg_launch_ms=stencil_ms / max(stencil_launches, 1),
        stencil_launches=int(stencil_launches),
        gather_ms_per_step=kern_ms[2] / steps,
        scatter_ms_per_step=kern_ms[3] / steps,
        stencil_gbs=B / (stencil_ms * 1e-3) / 1e9 if stencil_ms > 0 else None,
        lanczos=(lo.value, hi.value, it.value),
        stencil_config={"threads": thr.value, "points_div": pd.value, "z_chunks": zc.value,
                        "vector_runs": vr.value,
                        "tuner_ms": [round(tms[i], 4) for i in range(nt.value)]},
        finite=finite,
    )
    Xd.free(); Yd.free()
    return res


def cpu_baseline(sysm, X64, lam, lo, hi):
    """The oracle (single-threaded C restatement) on a bounded sample:
    the same grid, atoms and degree, 48 of the 160 columns."""
    ncol = 48
    op = sysm.oracle_op()
    Xs = np.asfortranarray(X64[:, :ncol])
    t0 = time.perf_counter()
    op.cheb_filter(Xs, sysm.veff, lam, lo, hi, DEGREE)
    dt = time.perf_counter() - t0
    B = filter_bytes(sysm.nd, ncol, DEGREE, 8)
    return {"value": B / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"oracle.c FP64 filter, degree {DEGREE}, {ncol} of {NCOL} columns on the same 60^3 Si64 "
                      f"grid with nonlocal projectors; {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--stencil-shape", default=None, metavar="VEC,THREADS,DIV,ZCHUNKS",
                    help="force the FP64 stencil launch shape instead of the autotuned one (e.g. 1,512,2,1)")
    args = ap.parse_args()
    shape = tuple(int(v) for v in args.stencil_shape.split(",")) if args.stencil_shape else None
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} (WORLD_SIZE {world}) is not run: the filter "
              "shards over k-points/structures/bands with no communication inside the filter (see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    from lynx_amd import _lib as L
    sysm = build_system()
    rng = np.random.default_rng(2)
    X64 = np.asfortranarray(rng.uniform(-0.5, 0.5, (sysm.nd, NCOL)))

    r64 = run_gpu(sysm, L.F64, args.steps, args.warmup, X64, shape)
    print(f"[bench] f64: {r64['ms_per_step']:.3f} ms/filter, {r64['gbs']:.0f} GB/s; stencil kernels "
          f"{r64['stencil_ms_per_step']:.3f} ms ({r64['stencil_gbs']:.0f} GB/s), gather "
          f"{r64['gather_ms_per_step']:.3f} ms, scatter {r64['scatter_ms_per_step']:.3f} ms", file=sys.stderr)
    variants = {}
    if not args.no_fp32:
        r32 = run_gpu(sysm, L.F32, args.steps, args.warmup, X64)
        print(f"[bench] f32: {r32['ms_per_step']:.3f} ms/filter, {r32['gbs']:.0f} GB/s", file=sys.stderr)
        variants["f32"] = {"value": round(r32["gbs"], 2), "ms_per_step": round(r32["ms_per_step"], 4),
                           "stencil_avg_launch_ms": round(r32["stencil_avg_launch_ms"], 5),
                           "roofline_frac": round(r32["stencil_gbs"] / HBM_PEAK_GBS, 4),
                           "speedup_vs_f64": round(r64["ms_per_step"] / r32["ms_per_step"], 3),
                           "launch_shape": r32["stencil_config"]}

    cpu = None
    if not args.no_cpu_baseline:
        lo, hi, _ = r64["lanczos"]
        cpu = cpu_baseline(sysm, X64, 0.5 * (lo + hi), lo, hi)

    achieved = r64["stencil_gbs"]
    kname = {0: "k_stencil_step<double", 1: "k_stencil_vec<double", 2: "k_stencil_vacc<double"}[
        r64["stencil_config"]["vector_runs"]]
    traffic, traffic_src = profiled_traffic(r64["stencil_config"], kname)
    line = {
        "metric": METRIC,
        "value": round(r64["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "Si64 filter-only operator bench: 2x2x2 diamond Si supercell a=20.52 Bohr, sigma=0.05 "
                        "Bohr (seed 1), 60^3 grid, FD order 12, 160 orbitals U[-0.5,0.5] (seed 2), Chebyshev "
                        "degree 20 over bounds from 10 Lanczos steps; C.pot psp8 stands in for Si; Veff Gaussian wells",
            "grid": [N, N, N], "ncol": NCOL, "degree": DEGREE, "fd_order": 12, "atoms": 64,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": {0: "k_stencil_step<double>", 1: "k_stencil_vec<double>",
                       2: "k_stencil_vacc<double>"}[r64["stencil_config"]["vector_runs"]]
            + " (fused stencil + Veff + Chebyshev recurrence)",
            "avg_launch_ms": round(r64["stencil_avg_launch_ms"], 5),
            "launch_shape": r64["stencil_config"],
        },
        "cpu_baseline": cpu,
        "variants": variants,
        "breakdown_ms_per_step": {"stencil": round(r64["stencil_ms_per_step"], 4),
                                  "nonlocal_gather": round(r64["gather_ms_per_step"], 4),
                                  "nonlocal_scatter": round(r64["scatter_ms_per_step"], 4)},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
