"""ctypes binding of liblynx_chefsi.so (the C ABI in include/lynx_chefsi.h).

The library is built in-tree by ``lynx_amd.build``; loading fails loudly when
it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblynx_chefsi.so")

OK, E_ARG, E_HIP, E_CHOLESKY, E_EIG, E_NONFINITE, E_IO, E_STATE, E_UNSUPPORTED = range(9)
F64, C128, F32, C64, BF16 = 0, 1, 2, 3, 4
FILTER_FP64, FILTER_FP32, FILTER_BF16 = 0, 1, 2
XC_LDA_PZ, XC_LDA_PW, XC_GGA_PBE = 0, 1, 2
SOLVE_SPECTRAL, SOLVE_AAR = 0, 1
GRAM_FP32, GRAM_BF16X3, GRAM_BF16X3_ROT = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1
CELL_ORTHOGONAL, CELL_NONORTHOGONAL = 0, 1

# every symbol include/lynx_chefsi.h declares, with (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_L = ctypes.c_long
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "lynx_ctx_create": (_I, [ctypes.POINTER(_P), _I]),
    "lynx_ctx_destroy": (None, [_P]),
    "lynx_last_error": (ctypes.c_char_p, [_P]),
    "lynx_status_string": (ctypes.c_char_p, [_I]),
    "lynx_ctx_stream": (_I, [_P, ctypes.POINTER(_P)]),
    "lynx_ctx_synchronize": (_I, [_P]),
    "lynx_set_grid": (_I, [_P, _I, _I, _I, _DP, _I, _I]),
    "lynx_get_grid_info": (_I, [_P, _DP, _DP, _DP]),
    "lynx_get_stencil": (_I, [_P, _DP, _DP, _DP]),
    "lynx_set_atoms": (_I, [_P, _I, ctypes.POINTER(ctypes.c_char_p), _I, _IP, _DP, _I]),
    "lynx_set_nonlocal": (_I, [_P, _I, _IP, _IP, _IP, _IP, _DP, _DP, _I, _IP, _DP]),
    "lynx_get_electrostatics": (_I, [_P, _DP, _DP, _DP, _DP, _DP, _DP]),
    "lynx_get_nonlocal_sizes": (_I, [_P, _IP, ctypes.POINTER(_L), ctypes.POINTER(_L), _IP, _IP]),
    "lynx_get_nonlocal": (_I, [_P, _IP, _IP, _IP, _IP, _DP, _DP, _IP, _DP]),
    "lynx_set_veff": (_I, [_P, _DP, _I]),
    "lynx_set_kpoint": (_I, [_P, _DP]),
    "lynx_h_apply": (_I, [_P, _I, _P, _I, _P, _I, _I, _D, _I]),
    "lynx_h_apply_noalloc": (_I, [_P, _I, _P, _I, _P, _I, _I, _D]),
    "lynx_cheb_filter": (_I, [_P, _I, _P, _I, _P, _I, _I, _I, _D, _D, _D, _I]),
    "lynx_orthogonalize": (_I, [_P, _I, _P, _I, _I, _I]),
    "lynx_project_hamiltonian": (_I, [_P, _I, _P, _I, _I, _P, _I]),
    "lynx_chefsi_solve": (_I, [_P, _I, _P, _I, _I, _DP, _D, _D, _D, _I, _DP, _I]),
    "lynx_lanczos_bounds": (_I, [_P, _I, _DP, _DP, _D, _I, ctypes.c_uint, _IP]),
    "lynx_fd_weights": (_I, [_I, _DP, _DP]),
    "lynx_randomize": (_I, [_P, _I, _P, _I, _I, _I, ctypes.c_uint, _I]),
    "lynx_psp8_info": (_I, [ctypes.c_char_p, _DP, _IP, _IP, _IP, _DP]),
    "lynx_kpoints": (_I, [_I, _I, _I, _DP, _DP, _I, _DP, _DP, _DP, _I, _IP]),
    "lynx_last_filter_timing": (_I, [_P, _DP, _IP]),
    "lynx_set_timing": (_I, [_P, _I]),
    "lynx_set_filter_graphs": (_I, [_P, _I]),
    "lynx_stream_ceiling": (_I, [_P, _L, _I, _DP]),
    "lynx_mixed_fallbacks": (_I, [_P, ctypes.POINTER(_L), ctypes.POINTER(_L)]),
    "lynx_set_filter_precision": (_I, [_P, _I]),
    "lynx_get_filter_precision": (_I, [_P, _IP]),
    "lynx_set_gram_precision": (_I, [_P, _I]),
    "lynx_get_gram_precision": (_I, [_P, _IP]),
    "lynx_gram": (_I, [_P, _I, _P, _I, _P, _I, _I, _P, _I]),
    "lynx_rotate": (_I, [_P, _I, _P, _I, _DP, _I, _P, _I, _I]),
    "lynx_get_veff_generation": (_I, [_P, ctypes.POINTER(_L)]),
    "lynx_solve_count": (_I, [_P, ctypes.POINTER(_L)]),
    "lynx_fermi_level": (_I, [_I, _I, _I, _DP, _DP, _D, _D, _I, _DP, _DP]),
    "lynx_get_grid_dims": (_I, [_P, _IP]),
    "lynx_scf_eig_create": (_I, [ctypes.POINTER(_P), _P, _P, _DP, _DP]),
    "lynx_scf_eig_destroy": (_I, [_P]),
    "lynx_scf_eig_last_error": (ctypes.c_char_p, [_P]),
    "lynx_scf_eig_step": (_I, [_P, _I, _D, _DP, ctypes.POINTER(_P), _I, _I, _DP, _DP, _DP, _IP]),
    "lynx_scf_eig_step_veff": (_I, [_P, _I, _D, _DP, _I, ctypes.POINTER(_P), _I, _I, _DP, _DP, _DP, _IP]),
    "lynx_scf_eig_state": (_I, [_P, _DP, _DP, _DP]),
    "lynx_scf_create": (_I, [ctypes.POINTER(_P), _P, _P, _DP, _DP]),
    "lynx_scf_destroy": (_I, [_P]),
    "lynx_scf_last_error": (ctypes.c_char_p, [_P]),
    "lynx_scf_set_initial_density": (_I, [_P, _DP, _I, _DP]),
    "lynx_scf_run": (_I, [_P, _P]),
    "lynx_scf_history": (_I, [_P, _I, _DP, _DP, _DP, _IP, _IP]),
    "lynx_scf_get": (_I, [_P, _DP, _DP, _DP, _DP]),
    "lynx_scf_get_fields": (_I, [_P, _DP, _DP, _DP]),
    "lynx_scf_build_potential": (_I, [_P, _DP]),
    "lynx_scf_forces": (_I, [_P, _DP, _DP]),
    "lynx_batch_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(_P), _DP]),
    "lynx_batch_destroy": (_I, [_P]),
    "lynx_batch_last_error": (ctypes.c_char_p, [_P]),
    "lynx_batch_filter": (_I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_L), ctypes.POINTER(_P),
                               ctypes.POINTER(_L), _IP, _I, _DP, _DP, _DP, _I]),
    "lynx_batch_set_graph": (_I, [_P, _I]),
    "lynx_batch_set_filter_precision": (_I, [_P, _I]),
    "lynx_batch_last_timing": (_I, [_P, _DP, _IP]),
    "lynx_batch_solve": (_I, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_L), _IP, _DP, _DP, _DP, _DP, _I, _DP]),
    "lynx_set_stencil_config": (_I, [_P, _I, _I, _I, _I]),
    "lynx_get_stencil_config": (_I, [_P, _I, _I, _IP, _IP, _IP, _IP, _DP, _IP]),
}

_lib = None


class LynxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(hipcc, gfx950). There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != OK:
        msg = lib().lynx_status_string(status).decode()
        if ctx:
            detail = lib().lynx_last_error(ctx)
            if detail:
                msg += ": " + detail.decode()
        raise LynxError(status, msg)


def dptr(a):
    return a.ctypes.data_as(_DP)


def iptr(a):
    return a.ctypes.data_as(_IP)


def vptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def dtype_code(arr_dtype):
    dt = np.dtype(arr_dtype)
    if dt == np.float64:
        return F64
    if dt == np.complex128:
        return C128
    if dt == np.float32:
        return F32
    if dt == np.complex64:
        return C64
    raise TypeError(f"unsupported dtype {dt}")


def np_dtype(code):
    return {F64: np.float64, C128: np.complex128, F32: np.float32, C64: np.complex64}[code]
