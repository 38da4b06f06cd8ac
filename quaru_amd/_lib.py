"""ctypes bindings of ``libquaru_hip.so`` (the C ABI in include/quaru_hip.h).

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing the import of the register fails with an
explicit error, and if no GPU is present the first register constructor
returns QS_ERR_NO_DEVICE.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libquaru_hip.so")

QS_C128, QS_C64 = 0, 1
QS_OK = 0
QS_INVALID_TARGET = 1
QS_INVALID_DIMENSIONS = 2
QS_INVALID_ARITY = 3
QS_INVALID_QUBIT = 4
QS_ERR_ARGUMENT = 16
QS_ERR_OUT_OF_MEMORY = 17
QS_ERR_HIP = 18
QS_ERR_NO_DEVICE = 19
QS_MEASURE_SLOTS = 64
# kernel classes of qs_timing_read (enum qs_kernel_class)
QS_K_SWEEP, QS_K_GATE1, QS_K_GATE2, QS_K_GATEK, QS_K_BIG, QS_K_DIAG, QS_K_SPARSE, QS_K_PERM = range(8)

# Every symbol include/quaru_hip.h declares, with (restype, argtypes).
_P = ctypes.c_void_p
_S = ctypes.c_size_t
_SP = ctypes.POINTER(ctypes.c_size_t)
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int
_IP = ctypes.POINTER(ctypes.c_int)
_U64 = ctypes.c_uint64
_U64P = ctypes.POINTER(ctypes.c_uint64)
_PP = ctypes.POINTER(ctypes.c_void_p)

SIGNATURES = {
    "qs_last_error": (_I, [_SP, _SP, ctypes.POINTER(ctypes.c_char_p)]),
    "qs_version": (ctypes.c_char_p, []),
    "qs_register_new": (_I, [_P, _S, _I, _PP]),
    "qs_register_from_int": (_I, [_S, _S, _I, _PP]),
    "qs_register_from_qubits": (_I, [_P, _P, _S, _I, _PP]),
    "qs_register_clone": (_I, [_P, _PP]),
    "qs_register_free": (None, [_P]),
    "qs_register_size": (_S, [_P]),
    "qs_register_dtype": (_I, [_P]),
    "qs_apply": (_I, [_P, _P, _S, _S, _P, _S]),
    "qs_apply_controlled": (_I, [_P, _P, _P, _S, _P, _P, _S]),
    "qs_apply_diagonal": (_I, [_P, _P, _P, _S]),
    "qs_apply_all": (_I, [_P, _P, _S, _S, _S]),
    "qs_measure": (_I, [_P, _S, _D, _IP]),
    "qs_measure_sequence": (_I, [_P, _P, _S, _P, _U64P]),
    "qs_sample_cdf": (_I, [_P, _P, _S, _P]),
    "qs_get_state": (_I, [_P, _P]),
    "qs_set_state": (_I, [_P, _P]),
    "qs_get_amplitudes": (_I, [_P, _P, _S, _P]),
    "qs_get_amplitudes_noalloc": (_I, [_P, _P, _S, _P]),
    "qs_norm_sqr": (_I, [_P, _DP]),
    "qs_probabilities": (_I, [_P, _P]),
    "qs_inner": (_I, [_P, _P, _DP, _DP]),
    "qs_equal": (_I, [_P, _P, _D, _IP]),
    "qs_sync": (_I, [_P]),
    "qs_set_fusion": (_I, [_P, _I]),
    "qs_stats": (_I, [_P, _U64P, _U64P, _U64P, _I]),
    "qs_measure_deferred": (_I, [_P, _S, _D, _S]),
    "qs_read_slots": (_I, [_P, _S, _S, _P]),
    "qs_stats_sweep_ops": (_I, [_P, _U64P, _U64P, _I]),
    "qs_timing_enable": (_I, [_P, _I]),
    "qs_timing_read": (_I, [_P, _P, _P, _I]),
    "qs_pcircuit_new": (_I, [_S, _PP]),
    "qs_pcircuit_free": (None, [_P]),
    "qs_pcircuit_gate": (_I, [_P, _P, _P, _S]),
    "qs_pcircuit_rotation": (_I, [_P, _I, _S, _S]),
    "qs_pcircuit_layer": (_I, [_P]),
    "qs_pcircuit_nparams": (_S, [_P]),
    "qs_adjoint_forward": (_I, [_P, _P, _S, _P, _P, _S, _I, _S, _PP, _P]),
    "qs_adjoint_backward": (_I, [_P, _P]),
    "qs_adjoint_stats": (_I, [_P, _P, _P]),
    "qs_adjoint_reverse_stats": (_I, [_P, _P, _P]),
    "qs_sweep_jit_mode": (_I, [_I]),
    "qs_sweep_jit_wait": (_I, []),
    "qs_sweep_jit_stats": (_I, [_P, _P, _P]),
    "qs_adjoint_free": (None, [_P]),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load the shared library once; raise LibraryMissing if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python -m quaru_amd.build` (needs hipcc)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    a, b = ctypes.c_size_t(0), ctypes.c_size_t(0)
    msg = ctypes.c_char_p()
    load().qs_last_error(ctypes.byref(a), ctypes.byref(b), ctypes.byref(msg))
    return a.value, b.value, (msg.value or b"").decode()
