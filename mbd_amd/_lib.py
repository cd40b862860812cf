"""ctypes binding of libmbd_hip.so (the C ABI declared in include/mbd_hip.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the library is missing the import of the engine fails loudly.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbd_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "mbd_hip.h")

MODE_FP64 = 0
MODE_MIXED = 1
MODE_LEARNED = 2

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


class MbdDiag(ctypes.Structure):
    _fields_ = [("info_scs", ctypes.c_int), ("fallback_used", ctypes.c_int),
                ("refine_iters", ctypes.c_int), ("n_negative", ctypes.c_int),
                ("info_eig", ctypes.c_int), ("supercell", ctypes.c_int * 3),
                ("residual", ctypes.c_double), ("t_scs_ms", ctypes.c_double),
                ("t_mbd_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double)]


class MbdBatchDiag(ctypes.Structure):
    _fields_ = [("config", ctypes.c_int), ("nb", ctypes.c_int), ("nt", ctypes.c_int),
                ("grid", ctypes.c_int), ("fallbacks", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("n_failed", ctypes.c_int), ("n_negative", ctypes.c_int),
                ("max_residual", ctypes.c_double), ("t_ms", ctypes.c_double),
                ("phase_ms", ctypes.c_double * 6)]


class LibraryMissing(RuntimeError):
    pass


_lib = None


def declared_symbols():
    """Function names declared in include/mbd_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(mbd_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                             "(the engine has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    h = ctypes.c_void_p
    L.mbd_create.argtypes = [ctypes.c_int, ctypes.POINTER(h)]
    L.mbd_destroy.argtypes = [h]
    L.mbd_last_error.argtypes = [h]
    L.mbd_last_error.restype = ctypes.c_char_p
    L.mbd_free_atom_params.argtypes = [ctypes.c_char_p, _D, _D, _D]
    L.mbd_beta.argtypes = [ctypes.c_int]
    L.mbd_beta.restype = ctypes.c_double
    L.mbd_casimir_grid.argtypes = [_D, _D]
    L.mbd_casimir_grid.restype = None
    L.mbd_set_system.argtypes = [h, ctypes.c_int, _D, _D, _D, _D, _D, _D, ctypes.c_int, _I,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double]
    L.mbd_scs.argtypes = [h, ctypes.c_int, _D, _D, _D, _D, _D, ctypes.POINTER(MbdDiag)]
    L.mbd_energy.argtypes = [h, _D, _D, _D, _D, _D, ctypes.POINTER(MbdDiag)]
    L.mbd_run.argtypes = [h, ctypes.c_int, _D, _D, _D, _D, _D, _D, ctypes.POINTER(MbdDiag)]
    _F = ctypes.POINTER(ctypes.c_float)
    L.mbd_net_layout.argtypes = [_I, _I, ctypes.POINTER(ctypes.c_long)]
    L.mbd_net_layout.restype = None
    L.mbd_net_set_weights.argtypes = [h, _F, ctypes.c_long]
    L.mbd_net_descriptors.argtypes = [h, _F, _D]
    L.mbd_net_forward.argtypes = [ctypes.c_int, _F, ctypes.c_int, _F, _F]
    L.mbd_net_grad.argtypes = [ctypes.c_int, _F, ctypes.c_int, _F, _F, _F, _F]
    L.mbd_net_train.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F, ctypes.c_int, ctypes.c_int,
                                ctypes.c_float, ctypes.c_uint, _F, _F]
    L.mbd_profile_enable.argtypes = [h, ctypes.c_int]
    L.mbd_profile_read.argtypes = [h, ctypes.c_int, _D, ctypes.POINTER(ctypes.c_char_p)]
    L.mbd_debug_relay_matrix.argtypes = [h, ctypes.c_int, ctypes.c_int, _D]
    L.mbd_debug_cfdm_matrix.argtypes = [h, _D, _D, _D, _D]
    L.mbd_debug_cholesky.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _I]
    L.mbd_debug_sym_eigvals.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
    L.mbd_debug_lu_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _I]
    L.mbd_debug_gemm_nt.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, _D, _D, ctypes.c_double, _D]
    L.mbd_batch_set.argtypes = [h, ctypes.c_int, _I, _D, _D, _D, _D, _D, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double]
    L.mbd_batch_tune.argtypes = [h, ctypes.c_int, ctypes.c_int, _D, _I]
    L.mbd_batch_set_config.argtypes = [h, ctypes.c_int]
    L.mbd_batch_config.argtypes = [ctypes.c_int, _I, _I, _I]
    L.mbd_batch_run.argtypes = [h, ctypes.c_int, _D, _D, _D, _D, _D, _I,
                                ctypes.POINTER(MbdBatchDiag)]
    L.mbd_debug_copy.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_double)]
    vp = ctypes.c_void_p
    L.mbd_batch_run_noalloc.argtypes = ([h, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int] + [vp] * 6 + [ctypes.c_double] * 3
                                        + [vp] * 5)
    for name in declared_symbols():
        getattr(L, name)  # every declared entry point must resolve
    _lib = L
    return L


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def dptr(a):
    return a.ctypes.data_as(_D)


def iptr(a):
    return a.ctypes.data_as(_I)
