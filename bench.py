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

Beside the headline line (under "variants"): the same filter in FP32; the
SCF wall-time of the full FP64 SCF on Si8 (30^3, PBE, 24 states) with the
oracle's CPU restatement timed beside it; Si216 (90^3, 480 states) eigen
stage and whole SCF in FP64, FP32-until-1e-4 and BF16-storage-until-1e-4
with their energy and eigenvalue errors against the FP64 run; the 112
k-point complex64 batch and the 64-frame heterogeneous batch.
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
    from systems import si64_filter_bench
    return si64_filter_bench()


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
    # per-kernel HIP events on the library's stream (lynx_set_timing); at
    # least two untimed calls: the first tunes the launch shape, the second
    # captures the filter's HIP graph (kernels and their event records) that
    # the timed calls replay
    L.check(lib.lynx_set_timing(ctx, 1), ctx)
    for _ in range(max(warmup, 2)):
        L.check(lib.lynx_cheb_filter(*args), ctx)
    kern_ms = np.zeros(4)
    kern_n = np.zeros(4, dtype=int)
    t_n = (ctypes.c_int * 4)()
    t_d = (ctypes.c_double * 4)()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        L.check(lib.lynx_cheb_filter(*args), ctx)
        L.check(lib.lynx_last_filter_timing(ctx, t_d, t_n), ctx)
        kern_ms += np.array(list(t_d))
        kern_n += np.array(list(t_n))
    synchronize()
    elapsed = time.perf_counter() - t0
    L.check(lib.lynx_set_timing(ctx, 0), ctx)
    thr = ctypes.c_int(); pd = ctypes.c_int(); zc = ctypes.c_int(); vr = ctypes.c_int(); nt = ctypes.c_int(128)
    tms = (ctypes.c_double * 128)()
    L.check(lib.lynx_get_stencil_config(ctx, dtype_code, NCOL, ctypes.byref(thr), ctypes.byref(pd),
                                        ctypes.byref(zc), ctypes.byref(vr), tms, ctypes.byref(nt)), ctx)
    Y = np.zeros((NCOL, nd), dtype=Xh.dtype)
    Yd.download(Y)
    finite = bool(np.all(np.isfinite(Y)))
    Y = Y.T      # (nd, ncol) view of the column-major block
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
        Y=Y,
    )
    Xd.free(); Yd.free()
    return res


def _bounds(ctx, lib, L, dtype_code, steps):
    lo = ctypes.c_double(); hi = ctypes.c_double(); it = ctypes.c_int()
    L.check(lib.lynx_lanczos_bounds(ctx, dtype_code, ctypes.byref(lo), ctypes.byref(hi), 0.0, steps, 1,
                                    ctypes.byref(it)), ctx)
    return lo.value, hi.value


def _batch_bytes(units, esize, vsize, degree):
    """Algorithmic bytes of one batched filter: s*P*(3m-1) plus one Veff read
    per unit and step."""
    return sum(esize * nd * nc * (3 * degree - 1) + vsize * nd * degree for nd, nc in units)


def _time_batch(bf, dtype, xs, ys, lds, ncols, lc, lo, hi, steps, warmup):
    from lynx_amd.hipmem import synchronize
    args = (dtype, [x.value for x in xs], [y.value for y in ys], lds, ncols, DEGREE, lc, lo, hi)
    for _ in range(warmup):
        bf.filter_device(*args)
    synchronize()
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        bf.filter_device(*args)
        dev.append(bf.last_timing()[0])
    synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps, float(np.median(dev)), bf.last_timing()[1]


def run_kpoint_batch(steps, warmup):
    """Batched k-points: Si2 FCC primitive cell (a = 10.26 Bohr), 24^3 grid,
    6x6x6 Monkhorst-Pack grid reduced by time reversal to 112 k-points,
    12 complex64 orbitals per k-point, all filtered in one graph-replayed
    launch per kernel and step; against the same filter per k-point in
    complex128 (the FP64 complex path) and in complex64, one call each.
    Then one whole CheFSI pass in complex128 (lynx_batch_solve against one
    lynx_chefsi_solve per k-point, best of two from the same start), and the
    batched pass with its filter in complex64."""
    from systems import si2_fcc
    from lynx_amd import _lib as L, core
    from lynx_amd.hipmem import DeviceArray, synchronize
    lib = L.lib()
    s = si2_fcc(24)
    H = s.hamiltonian(0)
    H.set_veff(s.veff)
    lv = np.ascontiguousarray(s.latvec, float)
    n = ctypes.c_int()
    sh = np.zeros(3)
    L.check(lib.lynx_kpoints(6, 6, 6, L.dptr(sh), L.dptr(lv), L.CELL_NONORTHOGONAL, None, None, None, 0,
                             ctypes.byref(n)))
    kc = np.zeros(3 * n.value)
    L.check(lib.lynx_kpoints(6, 6, 6, L.dptr(sh), L.dptr(lv), L.CELL_NONORTHOGONAL, L.dptr(kc), None, None,
                             n.value, ctypes.byref(n)))
    ks = kc.reshape(-1, 3)
    nk, nb, nd = len(ks), 12, s.nd
    H.set_kpoint(ks[0])
    lo, hi = _bounds(H.ctx, lib, L, L.C128, LANCZOS_STEPS)   # bounds at k-point 0 (SCF.cpp:883-935)
    lc = lo + 0.25 * (hi - lo)
    rng = np.random.default_rng(5)
    X = [np.asfortranarray(rng.uniform(-0.5, 0.5, (nd, nb)) + 1j * rng.uniform(-0.5, 0.5, (nd, nb)))
         for _ in range(nk)]
    x64 = [DeviceArray.from_numpy(x.T.astype(np.complex64)) for x in X]
    y64 = [DeviceArray(nd * nb * 8) for _ in range(nk)]
    bf = core.BatchFilter([(H, k) for k in ks])
    wall, dev, launches = _time_batch(bf, np.complex64, x64, y64, [nd] * nk, [nb] * nk, [lc] * nk, [lo] * nk,
                                      [hi] * nk, steps, warmup)
    B = _batch_bytes([(nd, nb)] * nk, 8, 4, DEGREE)
    # the same work one k-point per call (complex64, then complex128)
    per = {}
    for code, xs, es in ((L.C64, x64, 8), (L.C128, None, 16)):
        if xs is None:
            xs = [DeviceArray.from_numpy(x.T.copy()) for x in X]
        ys = [DeviceArray(nd * nb * es) for _ in range(nk)]
        for rep in range(2):
            synchronize()
            t0 = time.perf_counter()
            for i in range(nk):
                H.set_kpoint(ks[i])
                L.check(lib.lynx_cheb_filter(H.ctx, code, xs[i].value, nd, ys[i].value, nd, nb, DEGREE, lc, lo,
                                             hi, L.MEM_DEVICE), H.ctx)
            synchronize()
        per[code] = (1e3 * (time.perf_counter() - t0), ys, xs)
    # complex64 batch against the complex128 per-k-point filter
    errs = []
    for i in range(0, nk, 7):
        a = np.zeros((nb, nd), np.complex64); y64[i].download(a)
        b = np.zeros((nb, nd), np.complex128); per[L.C128][1][i].download(b)
        errs.append(float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    # a whole CheFSI pass over the 112 k-points in complex128: one batched
    # solve (filter, Cholesky-QR, H psi, Rayleigh-Ritz) against one
    # lynx_chefsi_solve per k-point, from the same start blocks
    solve = {}
    for mode in ("batched", "batched_fp32_filter", "per_kpoint"):
        bf.set_filter_precision("fp32" if mode == "batched_fp32_filter" else "fp64")
        blocks = [DeviceArray.from_numpy(x.T.copy()) for x in X]
        best, eig = 1e9, None
        for rep in range(2):
            for i, d in enumerate(blocks):
                d.upload(X[i].T.copy())
            synchronize()
            t0 = time.perf_counter()
            if mode.startswith("batched"):
                eig, _ = bf.solve_device(np.complex128, [d.value for d in blocks], [nd] * nk, [nb] * nk, DEGREE,
                                         [lc] * nk, [lo] * nk, [hi] * nk)
            else:
                eig = []
                ev = np.zeros(nb)
                for i in range(nk):
                    H.set_kpoint(ks[i])
                    L.check(lib.lynx_chefsi_solve(H.ctx, L.C128, blocks[i].value, nd, nb, L.dptr(ev), lc, lo, hi,
                                                  DEGREE, None, L.MEM_DEVICE), H.ctx)
                    eig.append(ev.copy())
            synchronize()
            best = min(best, time.perf_counter() - t0)
        solve[mode] = (1e3 * best, eig)
        for d in blocks:
            d.free()
    eig_diff = max(float(np.abs(a - b).max()) for a, b in zip(solve["batched"][1], solve["per_kpoint"][1]))
    eig_diff32 = max(float(np.abs(a - b).max())
                     for a, b in zip(solve["batched_fp32_filter"][1], solve["per_kpoint"][1]))
    bf.close()
    return {"workload": "Si2 FCC primitive cell a=10.26 Bohr, 24^3 grid, 6x6x6 MP -> %d k-points, %d complex64 "
                        "orbitals per k-point, degree %d, bounds from %d Lanczos steps at k-point 0"
                        % (nk, nb, DEGREE, LANCZOS_STEPS),
            "ms_per_filter": round(wall, 4), "device_ms_per_filter": round(dev, 4),
            "launches_per_filter": launches, "GBps": round(B / (dev * 1e-3) / 1e9, 1),
            "roofline_frac": round(B / (dev * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "per_kpoint_c64_ms": round(per[L.C64][0], 3), "per_kpoint_c128_ms": round(per[L.C128][0], 3),
            "speedup_vs_per_kpoint_c128": round(per[L.C128][0] / wall, 2),
            "max_rel_err_c64_batch_vs_c128": max(errs),
            "chefsi_pass_c128": {"batched_ms": round(solve["batched"][0], 3),
                                 "per_kpoint_ms": round(solve["per_kpoint"][0], 3),
                                 "speedup": round(solve["per_kpoint"][0] / solve["batched"][0], 2),
                                 "max_abs_ritz_value_diff_Ha": eig_diff,
                                 "batched_fp32_filter_ms": round(solve["batched_fp32_filter"][0], 3),
                                 "max_abs_ritz_value_diff_Ha_fp32_filter": eig_diff32}}


def run_hetero_batch(steps, warmup):
    """Heterogeneous batch: 32 displaced Si8 frames (30^3, 24 orbitals), 16
    Si16 frames (30x30x60, 40), 8 BaTiO3 frames (28^3, 30; La/Zr/O psp8
    stand-ins), 8 H2O frames (50^3, 8; O psp8, H in Veff only); sigma = 0.05
    Bohr (seed 4 + frame).  Filtered together in FP64 and FP32; per-class
    FP32-vs-FP64 error of the filtered blocks."""
    from systems import hetero_frames
    from lynx_amd import _lib as L, core
    from lynx_amd.hipmem import DeviceArray, synchronize
    lib = L.lib()
    frames = hetero_frames()
    t0 = time.perf_counter()
    Hs, bnds = [], []
    for _, s, nb in frames:
        H = s.hamiltonian(0)
        H.set_veff(s.veff)
        Hs.append(H)
        lo, hi = _bounds(H.ctx, lib, L, L.F64, LANCZOS_STEPS)
        bnds.append((lo + 0.25 * (hi - lo), lo, hi))
    setup_s = time.perf_counter() - t0
    rng = np.random.default_rng(6)
    X = [rng.uniform(-0.5, 0.5, (nb, s.nd)) for _, s, nb in frames]
    lc, lo, hi = (list(v) for v in zip(*bnds))
    res = {}
    outs = {}
    for dt, code, es in ((np.float64, L.F64, 8), (np.float32, L.F32, 4)):
        xs = [DeviceArray.from_numpy(x.astype(dt)) for x in X]
        ys = [DeviceArray(x.size * es) for x in X]
        bf = core.BatchFilter([(H, None) for H in Hs])
        wall, dev, launches = _time_batch(bf, dt, xs, ys, [s.nd for _, s, _ in frames], [nb for _, _, nb in frames],
                                          lc, lo, hi, steps, warmup)
        B = _batch_bytes([(s.nd, nb) for _, s, nb in frames], es, es, DEGREE)
        res[code] = {"ms_per_filter": round(wall, 4), "device_ms_per_filter": round(dev, 4),
                     "GBps": round(B / (dev * 1e-3) / 1e9, 1),
                     "roofline_frac": round(B / (dev * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "launches_per_filter": launches}
        outs[code] = [y.download(np.zeros(x.shape, dt)) for y, x in zip(ys, X)]
        bf.close()
        for d in xs + ys:
            d.free()
    # one frame per call (FP64), the unbatched baseline; the second sweep over
    # the frames is timed (the first pays each context's launch-shape tuning)
    xs = [DeviceArray.from_numpy(x) for x in X]
    ys = [DeviceArray(x.size * 8) for x in X]
    for rep in range(2):
        synchronize()
        t0 = time.perf_counter()
        for i, (_, s, nb) in enumerate(frames):
            L.check(lib.lynx_cheb_filter(Hs[i].ctx, L.F64, xs[i].value, s.nd, ys[i].value, s.nd, nb, DEGREE, lc[i],
                                         lo[i], hi[i], L.MEM_DEVICE), Hs[i].ctx)
        synchronize()
        unbatched = 1e3 * (time.perf_counter() - t0)
    err = {}
    for i, (cls, _, _) in enumerate(frames):
        a, b = outs[L.F32][i].astype(np.float64), outs[L.F64][i]
        err[cls] = max(err.get(cls, 0.0), float(np.linalg.norm(a - b) / np.linalg.norm(b)))
    P = sum(s.nd * nb for _, s, nb in frames)
    return {"workload": "32 Si8 (30^3, 24) + 16 Si16 (30x30x60, 40) + 8 BaTiO3 (28^3, 30) + 8 H2O (50^3, 8), "
                        "%d point-columns, degree %d" % (P, DEGREE),
            "f64": res[L.F64], "f32": res[L.F32], "unbatched_f64_ms": round(unbatched, 3),
            "speedup_f32_batch_vs_unbatched_f64": round(unbatched / res[L.F32]["ms_per_filter"], 2),
            "max_rel_err_f32_vs_f64": err, "setup_s": round(setup_s, 2)}


def run_mixed_si216():
    """Mixed precision on Si216: the 3x3x3 diamond supercell (30.78 Bohr),
    sigma = 0.03 Bohr (seed 3), 90^3 grid, 480 orbitals (seed 7), degree 20.
    The SCF eigen stage (lynx_scf_eig) runs iteration 0 (4 CheFSI passes from
    the random start, Lanczos bounds) and iteration 1 (one pass) four times
    from the same start: FP64 throughout; the FP32 pass (filter, Gram
    products and rotations in FP32, Cholesky and eigensolve in FP64) while
    the SCF residual exceeds 1e-4 (residual 1e-3 given for iteration 1); the
    same with the Gram products as split-BF16x3 matrix-core products; and
    that with the filter's blocks stored in bfloat16.  The band-energy
    difference per atom uses the occupations of the FP64 run (Fermi-Dirac,
    1000 K).  The first stage also pays the launch-shape tuning."""
    from systems import si216_mixed
    from lynx_amd import _lib as L, core
    from lynx_amd.hipmem import DeviceArray, synchronize
    s = si216_mixed()
    nb, ne, beta = 480, 4.0 * 216, 1.0 / (3.166811563e-6 * 1000.0)
    rng = np.random.default_rng(7)
    X0 = rng.uniform(-0.5, 0.5, (nb, s.nd))
    out = {}
    eigs = {}
    modes = (("fp64", 0.0, "fp32", "fp32"), ("fp32_until_1e-4", 1e-4, "fp32", "fp32"),
             ("fp32_bf16x3_gram_until_1e-4", 1e-4, "fp32", "bf16x3"), ("bf16_until_1e-4", 1e-4, "bf16", "bf16x3"))
    for name, thr, low, gram in modes:
        H = s.hamiltonian(0)
        drv = core.ScfEigenStage(H, nb, ne, beta, fp64_switch_residual=thr, low_precision=low, gram=gram)
        d = DeviceArray.from_numpy(X0)
        synchronize()
        t0 = time.perf_counter()
        e0, occ0, ef0, p0 = drv.step_device(0, 1.0, s.veff, [d.value], s.nd)
        synchronize()
        t1 = time.perf_counter()
        e1, occ1, ef1, p1 = drv.step_device(1, 1e-3, s.veff, [d.value], s.nd)
        synchronize()
        t2 = time.perf_counter()
        nfb, nq2 = ctypes.c_long(), ctypes.c_long()
        L.check(L.lib().lynx_mixed_fallbacks(H.ctx, ctypes.byref(nfb), ctypes.byref(nq2)), H.ctx)
        out[name] = {"iter0_s": round(t1 - t0, 3), "iter1_s": round(t2 - t1, 3), "filter_precision": [p0, p1],
                     "fermi": ef1, "fp32_passes_redone_in_fp64": nfb.value, "cholesky_qr2_passes": nq2.value}
        eigs[name] = (e1, occ1)
        d.free()
        del drv, H
    e64, f64 = eigs["fp64"]
    occ_n = int(ne // 2)
    res = {"workload": "Si216 3x3x3 supercell a=30.78 Bohr, sigma=0.03 (seed 3), 90^3 grid, 480 orbitals, degree 20; "
                       "SCF iteration 0 (4 passes + Lanczos) and 1 (1 pass) of the eigen stage",
           "fp64": out["fp64"], "mixed": out["fp32_until_1e-4"], "mixed_bf16x3_gram": out["fp32_bf16x3_gram_until_1e-4"],
           "mixed_bf16": out["bf16_until_1e-4"]}
    for key, name in (("", "fp32_until_1e-4"), ("_bf16x3_gram", "fp32_bf16x3_gram_until_1e-4"),
                      ("_bf16", "bf16_until_1e-4")):
        emx, _ = eigs[name]
        res["speedup_iter1" + key] = round(out["fp64"]["iter1_s"] / out[name]["iter1_s"], 3)
        res["speedup_iter0" + key] = round(out["fp64"]["iter0_s"] / out[name]["iter0_s"], 3)
        res["max_abs_eig_err_occupied_Ha" + key] = float(np.abs(emx[:occ_n] - e64[:occ_n]).max())
        res["band_energy_err_Ha_per_atom" + key] = float(abs((2.0 * f64 * (emx - e64)).sum()) / 216)
    return res


def run_si216_passes(reps=3):
    """Per-pass times of the Si216 eigen stage (the workload of
    run_mixed_si216): FP64; the reduced-precision pass with the FP32 filter
    and split-BF16x3 Gram products (FP32 rotations); the same with the
    rotations as split-BF16x3 products too.  Each mode runs SCF
    iteration 0 and one warm iteration 1, pauses 0.5 s (the pause separates
    the timed passes in a kernel trace), then times `reps` iteration-1
    passes."""
    from systems import si216_mixed
    from lynx_amd import core
    from lynx_amd.hipmem import DeviceArray, synchronize
    s = si216_mixed()
    nb, ne, beta = 480, 4.0 * 216, 1.0 / (3.166811563e-6 * 1000.0)
    X0 = np.random.default_rng(7).uniform(-0.5, 0.5, (nb, s.nd))
    out = {}
    for name, kw in (("fp64", {}),
                     ("fp32_bf16x3", dict(fp64_switch_residual=1e-4, low_precision="fp32", gram="bf16x3")),
                     ("fp32_bf16x3_rot", dict(fp64_switch_residual=1e-4, low_precision="fp32", gram="bf16x3_rot"))):
        H = s.hamiltonian(0)
        drv = core.ScfEigenStage(H, nb, ne, beta, **kw)
        d = DeviceArray.from_numpy(X0)
        drv.step_device(0, 1.0, s.veff, [d.value], s.nd)
        drv.step_device(1, 1e-3, s.veff, [d.value], s.nd)
        synchronize()
        time.sleep(0.5)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            drv.step_device(1, 1e-3, s.veff, [d.value], s.nd)
            synchronize()
            ts.append(time.perf_counter() - t0)
        out[name + "_pass_ms"] = round(1e3 * min(ts), 2)
        d.free()
        del drv, H
    out["speedup"] = round(out["fp64_pass_ms"] / out["fp32_bf16x3_pass_ms"], 3)
    out["speedup_rot"] = round(out["fp64_pass_ms"] / out["fp32_bf16x3_rot_pass_ms"], 3)
    return out


KB_HA = 3.1668115634556e-6   # Boltzmann constant, Ha/K


def _scf_history_summary(S):
    h = S.history()
    prec = [int(v) for v in h["precision"]]
    return {"iterations_per_precision": {str(p): prec.count(p) for p in sorted(set(prec))},
            "s_per_iteration": [round(float(t), 4) for t in h["seconds"]]}


def run_scf_si8(with_cpu):
    """Full FP64 SCF on the GPU: Si8 diamond cell (a = 10.26 Bohr, C.pot for
    Si), 30^3 grid, FD order 12, GGA-PBE, Fermi-Dirac 1000 K, 24 states,
    degree 20, orbitals from the glibc rand() stream, to a 1e-6 density
    residual.  One untimed run tunes the launch shapes; the timed run starts
    from the atomic density again.  With with_cpu the oracle's restatement of
    the same SCF (its filter in C, the rest numpy; one core) is timed beside
    it (both with the spectral Poisson and Kerker solves) and its converged
    energy is the check of the GPU run."""
    from systems import si8
    from lynx_amd.scf import SCF
    from lynx_amd.hipmem import synchronize
    s = si8(30)
    H = s.hamiltonian(0)
    beta = 1.0 / (KB_HA * 1000.0)
    kw = dict(xc="GGA_PBE", tol=1e-6)
    warm = SCF(H, 24, beta, **kw)
    warm.run()
    warm.close()
    S = SCF(H, 24, beta, **kw)
    synchronize()
    t0 = time.perf_counter()
    r = S.run()
    synchronize()
    wall = time.perf_counter() - t0
    out = {"workload": "Si8 a=10.26 Bohr, 30^3, FD12, GGA-PBE, Fermi-Dirac 1000 K, 24 states, degree 20, "
                       "SCF to 1e-6 density residual; C.pot stands in for Si",
           "scf_wall_s": round(wall, 4), "iterations": r["iterations"], "converged": bool(r["converged"]),
           "etotal_Ha": r["etotal"], "etotal_per_atom_Ha": r["etotal"] / 8}
    out.update(_scf_history_summary(S))
    S.close()
    if with_cpu:
        import oracle
        import oracle.scf as osc
        og = oracle.Grid(30, 30, 30, oracle.Lattice(np.diag([s.a] * 3), True))
        O = osc.ScfOracle(og, [oracle.load_psp8(f) for f in s.psp_files], [0] * 8, s.frac, 24, beta, xc=2,
                          tol=1e-6, solver="spectral")
        t0 = time.perf_counter()
        ref = O.run()
        dt = time.perf_counter() - t0
        out["cpu_restatement"] = {"scf_wall_s": round(dt, 3), "iterations": len(O.history), "cores": 1,
                                  "kind": "port", "cpu_model": _cpu_model(),
                                  "sample": "oracle SCF (CPU restatement, not the reference; filter in C, "
                                            "the rest numpy, spectral Poisson/Kerker) on the same input, one core"}
        out["etotal_err_per_atom_vs_oracle_Ha"] = abs(r["etotal"] - ref["etotal"]) / 8
        out["speedup_vs_cpu_restatement"] = round(dt / wall, 2)
    return out


def run_scf_si2_kpoints():
    """SCF wall-time with k-points: Si2 in the FCC primitive cell (a = 10.26
    Bohr, C.pot for Si), 24^3 grid, the 6x6x6 Monkhorst-Pack grid reduced by
    time reversal to 112 k-points, 12 states per k-point, GGA-PBE,
    Fermi-Dirac 1000 K, degree 20, orbitals from the glibc rand() stream, to
    a 1e-6 density residual; FP64 throughout, and with the filter in
    complex64 until the residual falls below 1e-4.  Each pass solves the
    112 k-points together (lynx_batch_solve).  One untimed iteration per
    mode tunes the launch shapes."""
    from systems import si2_fcc, fcc_kpoints
    from lynx_amd.scf import SCF
    from lynx_amd.hipmem import synchronize
    s = si2_fcc(24)
    H = s.hamiltonian(0)
    ks, w = fcc_kpoints(s)
    beta = 1.0 / (KB_HA * 1000.0)
    out = {"workload": "Si2 FCC primitive cell a=10.26 Bohr, 24^3, FD12, 6x6x6 MP -> %d k-points, 12 states, "
                       "GGA-PBE, Fermi-Dirac 1000 K, degree 20, SCF to 1e-6 density residual; C.pot stands "
                       "in for Si" % len(ks), "modes": {}}
    for name, thr in (("fp64", 0.0), ("fp32_until_1e-4", 1e-4)):
        kw = dict(xc="GGA_PBE", tol=1e-6, kpts=ks, kpt_weights=w, fp64_switch_residual=thr)
        warm = SCF(H, 12, beta, max_iter=1, **kw)
        warm.run()
        warm.close()
        S = SCF(H, 12, beta, **kw)
        synchronize()
        t0 = time.perf_counter()
        r = S.run()
        synchronize()
        wall = time.perf_counter() - t0
        m = {"scf_wall_s": round(wall, 4), "iterations": r["iterations"], "converged": bool(r["converged"]),
             "etotal_per_atom_Ha": r["etotal"] / 2}
        m.update(_scf_history_summary(S))
        S.close()
        out["modes"][name] = m
    a, b = out["modes"]["fp64"], out["modes"]["fp32_until_1e-4"]
    out["speedup_scf_wall_fp32_until_1e-4"] = round(a["scf_wall_s"] / b["scf_wall_s"], 3)
    out["etotal_err_per_atom_Ha_fp32_until_1e-4"] = abs(a["etotal_per_atom_Ha"] - b["etotal_per_atom_Ha"])
    return out


def run_scf_si216():
    """SCF wall-time on Si216 (the 3x3x3 supercell, a = 30.78 Bohr,
    sigma = 0.03 Bohr seed 3, 90^3 grid, 480 states, GGA-PBE at Gamma,
    Fermi-Dirac 1000 K, degree 20, to a 1e-6 density residual) in four
    modes: FP64 throughout; FP32 passes (filter, Gram products, rotation)
    until the residual falls below 1e-4, FP64 after; the same with
    split-BF16x3 Gram products; and that with the filter's blocks stored in
    bfloat16.  Each mode first runs one untimed
    iteration so the launch shapes it uses are tuned.  Errors are against
    the FP64 run: total energy per atom, occupied eigenvalues and atomic
    forces (lynx_scf_forces after each run, timed)."""
    from systems import si216_mixed
    from lynx_amd.scf import SCF
    from lynx_amd.hipmem import synchronize
    s = si216_mixed()
    H = s.hamiltonian(0)
    beta = 1.0 / (KB_HA * 1000.0)
    modes = (("fp64", 0.0, "fp32", "fp32"), ("fp32_until_1e-4", 1e-4, "fp32", "fp32"),
             ("fp32_bf16x3_gram_until_1e-4", 1e-4, "fp32", "bf16x3"), ("bf16_until_1e-4", 1e-4, "bf16", "bf16x3"))
    out, res = {}, {}
    for name, thr, low, gram in modes:
        kw = dict(xc="GGA_PBE", tol=1e-6, fp64_switch_residual=thr, low_precision=low, gram=gram)
        warm = SCF(H, 480, beta, max_iter=1, **kw)
        warm.run()
        warm.close()
        S = SCF(H, 480, beta, **kw)
        synchronize()
        t0 = time.perf_counter()
        r = S.run()
        synchronize()
        wall = time.perf_counter() - t0
        eig, occ, _, _ = S.state()
        t0 = time.perf_counter()
        F = S.forces()
        tf = time.perf_counter() - t0
        res[name] = (r, eig, occ, F)
        out[name] = {"scf_wall_s": round(wall, 3), "iterations": r["iterations"], "converged": bool(r["converged"]),
                     "etotal_per_atom_Ha": r["etotal"] / 216, "forces_s": round(tf, 3),
                     "max_abs_force_Ha_per_Bohr": float(np.abs(F).max())}
        out[name].update(_scf_history_summary(S))
        S.close()
        print(f"[bench] Si216 SCF {name}: {wall:.2f} s, {r['iterations']} iterations, "
              f"E/atom {r['etotal'] / 216:.10f}", file=sys.stderr, flush=True)
    r64, e64, o64, F64 = res["fp64"]
    occ_mask = o64 > 1e-6
    summary = {"workload": "Si216 3x3x3 supercell a=30.78 Bohr, sigma=0.03 (seed 3), 90^3, GGA-PBE, Gamma, "
                           "480 states, Fermi-Dirac 1000 K, degree 20, SCF to 1e-6 density residual; C.pot "
                           "stands in for Si", "modes": out}
    for name in ("fp32_until_1e-4", "fp32_bf16x3_gram_until_1e-4", "bf16_until_1e-4"):
        r, e, _, F = res[name]
        summary[name] = {"speedup_scf_wall": round(out["fp64"]["scf_wall_s"] / out[name]["scf_wall_s"], 3),
                         "etotal_err_per_atom_Ha": abs(r["etotal"] - r64["etotal"]) / 216,
                         "max_abs_occupied_eig_err_Ha": float(np.abs(e - e64)[occ_mask].max()),
                         "max_force_err_Ha_per_Bohr": float(np.abs(F - F64).max())}
    return summary


BF16_MFMA_PEAK_TFLOPS, FP32_PEAK_TFLOPS, FP64_PEAK_TFLOPS = 2500.0, 157.3, 78.6


def run_gram_si216(reps=5):
    """The overlap product dV X^T X of the Si216 pass (729,000 grid points x
    480 orbitals, uniform random blocks resident in HBM) through lynx_gram:
    FP64 (rocBLAS split-K dgemm), FP32 (rocBLAS split-K sgemm) and
    split-BF16x3 (the hand-written matrix-core kernel).  Best of reps calls
    after one untimed call; each call includes the 1.8 MB result copy to
    the host.  Effective rate 2 K N^2 / t; for split-BF16x3 also the rate of
    its three bfloat16 products against the dense bf16 peak."""
    from lynx_amd import _lib as L
    from lynx_amd.hipmem import DeviceArray
    from systems import si216_mixed
    s = si216_mixed()
    H = s.hamiltonian(0)
    nd, n = s.nd, 480
    ld = ((nd + 63) // 64) * 64
    rng = np.random.default_rng(11)
    X = rng.uniform(-0.5, 0.5, (n, ld))
    d64 = DeviceArray.from_numpy(X)
    d32 = DeviceArray.from_numpy(X.astype(np.float32))
    C = np.zeros((n, n))
    out = {}
    flops = 2.0 * nd * n * n
    ref = None
    for name, dt, mode, buf in (("fp64", L.F64, L.GRAM_FP32, d64), ("fp32_gemm", L.F32, L.GRAM_FP32, d32),
                                ("split_bf16x3", L.F32, L.GRAM_BF16X3, d32)):
        L.check(L.lib().lynx_set_gram_precision(H.ctx, mode), H.ctx)
        best = 1e9
        for i in range(reps + 1):
            t0 = time.perf_counter()
            L.check(L.lib().lynx_gram(H.ctx, dt, buf.value, ld, buf.value, ld, n, L.dptr(C), L.MEM_DEVICE), H.ctx)
            dt_s = time.perf_counter() - t0
            if i > 0:
                best = min(best, dt_s)
        r = {"ms": round(best * 1e3, 3), "effective_tflops": round(flops / best / 1e12, 1)}
        if name == "fp64":
            ref = C.copy()
            r["frac_of_fp64_peak"] = round(flops / best / 1e12 / FP64_PEAK_TFLOPS, 3)
        else:
            r["max_rel_err_vs_fp64"] = float(np.abs(C - ref).max() / np.abs(ref).max())
        if name == "fp32_gemm":
            r["frac_of_fp32_peak"] = round(flops / best / 1e12 / FP32_PEAK_TFLOPS, 3)
        if name == "split_bf16x3":
            r["bf16_mfma_tflops"] = round(3 * flops / best / 1e12, 1)
            r["frac_of_bf16_mfma_peak"] = round(3 * flops / best / 1e12 / BF16_MFMA_PEAK_TFLOPS, 3)
        out[name] = r
    d64.free(); d32.free()
    out["workload"] = "dV X^T X, X 729,000 x 480 (Si216 90^3 grid, 480 orbitals), device-resident"
    return out


def stream_ceiling(sysm):
    """Practical HBM ceiling of the stencil step's traffic on this device
    (lynx_stream_ceiling: x, Y_(k-2) and y streams of the same 60^3 x 160
    FP64 shape plus an L2-resident Veff, no stencil), in the algorithmic
    bytes of the step per second."""
    from lynx_amd import _lib as L
    H = sysm.hamiltonian(0)
    g = ctypes.c_double()
    L.check(L.lib().lynx_stream_ceiling(H.ctx, sysm.nd, NCOL, ctypes.byref(g)), H.ctx)
    return round(g.value, 1)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sysm, X64, lam, lo, hi, Y_gpu):
    """The oracle's C restatement of the filter on the same workload (same
    grid, atoms, degree, bounds and all 160 columns), its columns shared among
    OpenMP threads on the host's cores (OMP_NUM_THREADS, else every core).
    Its output is also the check of the GPU run: the largest per-column
    relative difference over all 160 columns."""
    ncol = NCOL
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    op = sysm.oracle_op()
    Xs = np.asfortranarray(X64[:, :ncol])
    t0 = time.perf_counter()
    ref = op.cheb_filter(Xs, sysm.veff, lam, lo, hi, DEGREE, threads=threads)
    dt = time.perf_counter() - t0
    B = filter_bytes(sysm.nd, ncol, DEGREE, 8)
    err = float((np.linalg.norm(Y_gpu - ref, axis=0) / np.linalg.norm(ref, axis=0)).max())
    cpu = {"value": B / dt / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
           "cpu_model": _cpu_model(),
           "sample": f"oracle.c FP64 filter (CPU restatement, not the reference), degree {DEGREE}, all {ncol} "
                     f"columns on the same 60^3 Si64 grid with nonlocal projectors, one call on {threads} "
                     f"OpenMP threads; {dt:.2f} s"}
    return cpu, err


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batches", action="store_true", help="skip the k-point and heterogeneous batch legs")
    ap.add_argument("--no-si216", action="store_true", help="skip the Si216 mixed-precision legs")
    ap.add_argument("--no-scf", action="store_true", help="skip the SCF wall-time legs (Si8 and Si216)")
    ap.add_argument("--stencil-shape", default=None, metavar="VEC,THREADS,DIV,ZCHUNKS",
                    help="force the FP64 stencil launch shape instead of the autotuned one (e.g. 1,512,2,1)")
    ap.add_argument("--si216-passes", action="store_true",
                    help="only time single Si216 CheFSI passes (FP64 and the mixed pass) and print them")
    args = ap.parse_args()
    if args.si216_passes:
        print(json.dumps({"si216_passes": run_si216_passes()}), flush=True)
        return
    shape = tuple(int(v) for v in args.stencil_shape.split(",")) if args.stencil_shape else None
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} (WORLD_SIZE {world}) is not run: the filter "
              "shards over k-points/structures/bands with no communication inside the filter",
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

    if not args.no_scf:
        variants["scf_si8"] = run_scf_si8(not args.no_cpu_baseline)
        print(f"[bench] Si8 SCF: {variants['scf_si8']}", file=sys.stderr, flush=True)
        variants["scf_si2_kpoints"] = run_scf_si2_kpoints()
        print(f"[bench] Si2 k-point SCF: {variants['scf_si2_kpoints']}", file=sys.stderr, flush=True)
    if not args.no_si216:
        variants["gram_si216"] = run_gram_si216()
        print(f"[bench] Si216 Gram: {variants['gram_si216']}", file=sys.stderr, flush=True)
        variants["si216_mixed_precision"] = run_mixed_si216()
        print(f"[bench] Si216 mixed: {variants['si216_mixed_precision']}", file=sys.stderr)
        if not args.no_scf:
            variants["scf_si216"] = run_scf_si216()
            print(f"[bench] Si216 SCF: {variants['scf_si216']}", file=sys.stderr, flush=True)

    if not args.no_batches:
        variants["c64_kpoint_batch"] = run_kpoint_batch(args.steps, args.warmup)
        print(f"[bench] k-point batch: {variants['c64_kpoint_batch']}", file=sys.stderr)
        variants["hetero_batch"] = run_hetero_batch(args.steps, args.warmup)
        print(f"[bench] hetero batch: {variants['hetero_batch']}", file=sys.stderr)

    cpu = None
    parity = {"finite": r64["finite"], "max_col_rel_err_vs_oracle": None}
    if not args.no_cpu_baseline:
        lo, hi, _ = r64["lanczos"]
        cpu, err = cpu_baseline(sysm, X64, 0.5 * (lo + hi), lo, hi, r64["Y"])
        parity.update(max_col_rel_err_vs_oracle=err, columns_checked=NCOL, tolerance=1e-11)
    r64.pop("Y", None)

    achieved = r64["stencil_gbs"]
    ceiling = stream_ceiling(sysm)
    kname = {0: "k_stencil_step<double", 1: "k_stencil_vec<double", 2: "k_stencil_vacc<double",
             3: "k_stencil_vacc<double"}[3 &
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
            "kernel": {0: "k_stencil_step<double>", 1: "k_stencil_vec<double>", 2: "k_stencil_vacc<double>",
                       3: "k_stencil_vacc<double> (2 planes in flight)"}[r64["stencil_config"]["vector_runs"] & 3]
            + " (fused stencil + Veff + Chebyshev recurrence)",
            "avg_launch_ms": round(r64["stencil_avg_launch_ms"], 5),
            "launch_shape": r64["stencil_config"],
            "stream_ceiling": ceiling,
            "frac_of_stream_ceiling": round(achieved / ceiling, 4),
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "variants": variants,
        "breakdown_ms_per_step": {"stencil": round(r64["stencil_ms_per_step"], 4),
                                  "nonlocal_gather": round(r64["gather_ms_per_step"], 4),
                                  "nonlocal_scatter": round(r64["scatter_ms_per_step"], 4)},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
