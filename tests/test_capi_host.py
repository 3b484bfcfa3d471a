"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-side set-up (FD weights, stencil coefficients,
psp8 reader, nonlocal projectors) matches the oracle on the same inputs."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from systems import A_SI, DIAMOND, ROOT, displace, psp, supercell
from lynx_amd import _lib as L


def header_symbols():
    with open(os.path.join(ROOT, "include", "lynx_chefsi.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(lynx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES)


def host_ctx():
    p = ctypes.c_void_p()
    L.check(L.lib().lynx_ctx_create(ctypes.byref(p), -1))
    return p


def test_fd_weights_match_oracle():
    for order in (2, 6, 12, 24):
        n = order // 2 + 1
        w1 = np.zeros(n); w2 = np.zeros(n)
        L.check(L.lib().lynx_fd_weights(order, L.dptr(w1), L.dptr(w2)))
        r1, r2 = oracle.fd_weights(order)
        assert np.array_equal(w1, r1) and np.array_equal(w2, r2), order
    assert L.lib().lynx_fd_weights(7, None, None) == L.E_ARG


@pytest.mark.parametrize("latvec,cell", [
    (np.diag([10.26, 9.0, 12.0]), L.CELL_ORTHOGONAL),
    (0.5 * A_SI * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]]), L.CELL_NONORTHOGONAL),
])
def test_stencil_and_grid_info_match_oracle(latvec, cell):
    c = host_ctx()
    lv = np.ascontiguousarray(latvec, float)
    L.check(L.lib().lynx_set_grid(c, 24, 30, 36, L.dptr(lv), cell, 12), c)
    D1 = np.zeros(21); D2 = np.zeros(21); D2m = np.zeros(21)
    L.check(L.lib().lynx_get_stencil(c, L.dptr(D1), L.dptr(D2), L.dptr(D2m)), c)
    h = np.zeros(3); dV = ctypes.c_double(); me = ctypes.c_double()
    L.check(L.lib().lynx_get_grid_info(c, L.dptr(h), ctypes.byref(dV), ctypes.byref(me)), c)
    lat = oracle.Lattice(lv, cell == L.CELL_ORTHOGONAL)
    g = oracle.Grid(24, 30, 36, lat)
    st = oracle.stencil_coeffs(12, g)
    assert np.allclose(D1.reshape(3, 7), st.D1, rtol=1e-15, atol=0)
    assert np.allclose(D2.reshape(3, 7), st.D2, rtol=1e-14, atol=0)
    assert np.allclose(D2m.reshape(3, 7), st.D2m, rtol=1e-14, atol=0)
    assert np.allclose(h, g.h, rtol=1e-15) and abs(dV.value - g.dV) < 1e-15 * g.dV
    assert abs(me.value - st.max_eig) < 1e-12 * st.max_eig
    L.lib().lynx_ctx_destroy(c)


def test_psp8_info():
    z = ctypes.c_double(); lm = ctypes.c_int(); ll = ctypes.c_int(); npj = ctypes.c_int(); rc = ctypes.c_double()
    L.check(L.lib().lynx_psp8_info(psp("C.pot").encode(), ctypes.byref(z), ctypes.byref(lm), ctypes.byref(ll),
                                   ctypes.byref(npj), ctypes.byref(rc)))
    p = oracle.load_psp8(psp("C.pot"))
    assert (z.value, lm.value, ll.value, npj.value) == (4.0, 1, 4, 8)
    assert rc.value == p.rc_max
    assert L.lib().lynx_psp8_info(b"/nonexistent/file.psp8", None, None, None, None, None) == L.E_IO


def get_nonlocal(c):
    ni = ctypes.c_int(); ng = ctypes.c_long(); nc = ctypes.c_long(); na = ctypes.c_int(); npj = ctypes.c_int()
    L.check(L.lib().lynx_get_nonlocal_sizes(c, ctypes.byref(ni), ctypes.byref(ng), ctypes.byref(nc),
                                            ctypes.byref(na), ctypes.byref(npj)), c)
    out = dict(img_atom=np.zeros(ni.value, np.int32), img_ndc=np.zeros(ni.value, np.int32),
               img_nproj=np.zeros(ni.value, np.int32), gpos=np.zeros(ng.value, np.int32),
               chi=np.zeros(nc.value), img_shift=np.zeros(3 * ni.value), ip_displ=np.zeros(na.value + 1, np.int32),
               gamma=np.zeros(npj.value))
    L.check(L.lib().lynx_get_nonlocal(c, L.iptr(out["img_atom"]), L.iptr(out["img_ndc"]), L.iptr(out["img_nproj"]),
                                      L.iptr(out["gpos"]), L.dptr(out["chi"]), L.dptr(out["img_shift"]),
                                      L.iptr(out["ip_displ"]), L.dptr(out["gamma"])), c)
    return out


@pytest.mark.parametrize("n,a,frac", [
    (24, A_SI, DIAMOND),
    (30, 2 * A_SI, displace(supercell(2), 2 * A_SI, 0.05, seed=1)),
])
def test_nonlocal_builder_matches_oracle(n, a, frac):
    c = host_ctx()
    lv = np.diag([a, a, a]).astype(float)
    L.check(L.lib().lynx_set_grid(c, n, n, n, L.dptr(lv), L.CELL_ORTHOGONAL, 12), c)
    paths = (ctypes.c_char_p * 1)(psp("C.pot").encode())
    types = np.zeros(len(frac), np.int32)
    coords = np.ascontiguousarray(frac, float)
    L.check(L.lib().lynx_set_atoms(c, 1, paths, len(types), L.iptr(types), L.dptr(coords), 1), c)
    got = get_nonlocal(c)
    lat = oracle.Lattice(lv, True)
    g = oracle.Grid(n, n, n, lat)
    ps = [oracle.load_psp8(psp("C.pot"))]
    op = oracle.setup_system(g, ps, list(types), frac, 12)
    assert len(got["img_atom"]) == len(op.images)
    assert np.array_equal(got["img_atom"], [im.atom for im in op.images])
    assert np.array_equal(got["img_ndc"], [len(im.gpos) for im in op.images])
    assert np.array_equal(got["gpos"], np.concatenate([im.gpos for im in op.images]))
    assert np.allclose(got["img_shift"], np.concatenate([im.shift for im in op.images]), rtol=0, atol=1e-12)
    chi = np.concatenate([np.asfortranarray(im.chi).ravel(order="F") for im in op.images])
    assert np.abs(got["chi"] - chi).max() < 1e-13 * np.abs(chi).max()
    ip, gam = oracle.atom_gamma(ps, list(types))
    assert np.array_equal(got["ip_displ"], ip) and np.array_equal(got["gamma"], gam)
    L.lib().lynx_ctx_destroy(c)


def test_host_context_errors():
    lib = L.lib()
    c = host_ctx()
    lv = np.diag([A_SI] * 3).astype(float)
    x = np.zeros(8)
    assert lib.lynx_h_apply(c, L.F64, L.vptr(x), 0, L.vptr(x), 0, 1, 0.0, L.MEM_HOST) == L.E_STATE
    assert lib.lynx_set_grid(c, 0, 4, 4, L.dptr(lv), 0, 12) == L.E_ARG
    assert lib.lynx_set_grid(c, 8, 8, 8, L.dptr(lv), 0, 5) == L.E_ARG
    L.check(lib.lynx_set_grid(c, 8, 8, 8, L.dptr(lv), 0, 12), c)
    paths = (ctypes.c_char_p * 2)(psp("C.pot").encode(), b"/missing.psp8")
    types = np.array([0, 1], np.int32)
    coords = np.zeros(6)
    assert lib.lynx_set_atoms(c, 2, paths, 2, L.iptr(types), L.dptr(coords), 1) == L.E_IO
    assert b"missing" in lib.lynx_last_error(c)
    types = np.array([1, 0], np.int32)
    paths = (ctypes.c_char_p * 2)(psp("C.pot").encode(), psp("C.pot").encode())
    assert lib.lynx_set_atoms(c, 2, paths, 2, L.iptr(types), L.dptr(coords), 1) == L.E_ARG
    v = np.zeros(512)
    assert lib.lynx_set_veff(c, L.dptr(v), L.MEM_HOST) == L.E_STATE
    lib.lynx_ctx_destroy(c)


def test_gpu_context_without_device_fails_cleanly():
    # on a host with no GPU the call returns an error instead of aborting
    import subprocess
    import sys
    code = ("import ctypes,sys; sys.path.insert(0, %r); from lynx_amd import _lib as L; "
            "p=ctypes.c_void_p(); s=L.lib().lynx_ctx_create(ctypes.byref(p), 99); print(s)") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert int(r.stdout.strip().splitlines()[-1]) in (L.E_ARG, L.E_HIP)


def test_cpp_header_builds_and_runs_host_calls(tmp_path):
    # the C++ drop-in classes compile with g++, link the library, and run
    # the host-side set-up through a host-only context
    import subprocess
    exe = tmp_path / "host_api"
    lib_dir = os.path.join(ROOT, "lynx_amd")
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "host_api.cpp"), "-L", lib_dir, "-llynx_chefsi",
           f"-Wl,-rpath,{lib_dir}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe), psp("C.pot")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    nimg, ngpos, natom, nproj, threw = map(int, out.stdout.split())
    assert (natom, nproj, threw) == (8, 64, 1) and nimg >= 8 and ngpos > 8 * 100


@pytest.mark.parametrize("K,shift,latvec,cell", [
    ((6, 6, 6), (0.0, 0.0, 0.0), 0.5 * A_SI * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]]), L.CELL_NONORTHOGONAL),
    ((4, 3, 2), (0.5, 0.5, 0.0), np.diag([10.0, 11.0, 12.0]), L.CELL_ORTHOGONAL),
    ((3, 3, 3), (0.0, 0.5, 0.0), np.diag([7.0, 7.0, 9.0]), L.CELL_ORTHOGONAL),
])
def test_kpoints_match_oracle(K, shift, latvec, cell):
    lv = np.ascontiguousarray(latvec, float)
    sh = np.asarray(shift, float)
    n = ctypes.c_int()
    L.check(L.lib().lynx_kpoints(*K, L.dptr(sh), L.dptr(lv), cell, None, None, None, 0, ctypes.byref(n)))
    kc = np.zeros(3 * n.value); kr = np.zeros(3 * n.value); w = np.zeros(n.value)
    L.check(L.lib().lynx_kpoints(*K, L.dptr(sh), L.dptr(lv), cell, L.dptr(kc), L.dptr(kr), L.dptr(w), n.value,
                                 ctypes.byref(n)))
    rc, rr, rw = oracle.kpoints(K, shift, oracle.Lattice(lv, cell == L.CELL_ORTHOGONAL))
    assert n.value == len(rw)
    assert np.array_equal(kc.reshape(-1, 3), rc) and np.array_equal(kr.reshape(-1, 3), rr)
    assert np.array_equal(w, rw) and abs(w.sum() - 1) < 1e-12
    if K == (6, 6, 6):
        assert n.value == 112


def test_nonlocal_builder_multitype_matches_oracle():
    # three species, f projectors (La.pot, lmax 3) and a small cell
    from systems import batio3
    sysm = batio3(0, n=20)
    c = host_ctx()
    lv = np.ascontiguousarray(sysm.latvec, float)
    L.check(L.lib().lynx_set_grid(c, 20, 20, 20, L.dptr(lv), L.CELL_ORTHOGONAL, 12), c)
    paths = (ctypes.c_char_p * 3)(*[p.encode() for p in sysm.psp_files])
    coords = np.ascontiguousarray(sysm.frac, float)
    L.check(L.lib().lynx_set_atoms(c, 3, paths, len(sysm.types), L.iptr(sysm.types), L.dptr(coords), 1), c)
    got = get_nonlocal(c)
    op = sysm.oracle_op()
    assert np.array_equal(got["img_atom"], [im.atom for im in op.images])
    assert np.array_equal(got["gpos"], np.concatenate([im.gpos for im in op.images]))
    chi = np.concatenate([np.asfortranarray(im.chi).ravel(order="F") for im in op.images])
    assert np.abs(got["chi"] - chi).max() < 1e-12 * np.abs(chi).max()
    ps = [oracle.load_psp8(f) for f in sysm.psp_files]
    ip, gam = oracle.atom_gamma(ps, list(sysm.types))
    assert np.array_equal(got["ip_displ"], ip) and np.array_equal(got["gamma"], gam)
    assert ip[1] - ip[0] == 32    # La: 2 projectors for each of l = 0..3
    L.lib().lynx_ctx_destroy(c)
