"""CPU FP64 oracle for quaru's state-vector hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline), never as the thing measured or shipped.
The product library (``quaru_amd``) never imports it.

Thin ctypes bindings over ``liborc.so`` (built from ``qs_oracle.c`` and
``qs_oracle_omp.c`` by ``make -C oracle``).  Each binding follows the C
function of the same name, which cites the reference file:line it restates:

* :func:`new`              -- ``Register::new`` (src/register.rs:41-56)
* :func:`from_int`         -- ``Register::from_int`` (src/register.rs:58-63)
* :func:`try_from_qubits`  -- ``Register::try_from_qubits`` (src/register.rs:81-126)
* :func:`try_apply_fast`   -- ``Register::try_apply_fast`` (src/register.rs:140-191)
* :func:`try_apply`        -- ``Register::try_apply`` (src/register.rs:193-250)
* :func:`try_apply_all`    -- ``Register::try_apply_all`` (src/register.rs:267-281)
* :func:`try_measure`      -- ``Register::try_measure`` (src/register.rs:297-342)

Errors come back as ``(code, a, b)`` with the codes of ``OperationError``
(src/register.rs:10-20): 1 InvalidTarget(a), 2 InvalidDimensions(a, b),
3 InvalidArity(a), 4 InvalidQubit(a).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")

OK, INVALID_TARGET, INVALID_DIMENSIONS, INVALID_ARITY, INVALID_QUBIT = 0, 1, 2, 3, 4

_lib = None


def build() -> str:
    """Compile liborc.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        SP = ctypes.POINTER(ctypes.c_size_t)
        L.orc_operation_valid.argtypes = [S, S, S]
        L.orc_new.argtypes = [P, S, P]
        L.orc_from_int.argtypes = [S, S, P]
        L.orc_try_from_qubits.argtypes = [P, P, S, P, SP]
        L.orc_try_apply_fast.argtypes = [P, S, P, S, S, P, S, SP, SP]
        L.orc_try_apply.argtypes = [P, S, P, S, S, P, S, SP, SP]
        L.orc_try_apply_all.argtypes = [P, S, P, S, SP]
        L.orc_try_measure.argtypes = [P, S, S, ctypes.c_double, ctypes.POINTER(ctypes.c_int), SP]
        L.orc_measure_sequence.argtypes = [P, S, P, S, P]
        L.orc_measure_sequence.restype = S
        L.orc_apply_fast_omp.argtypes = [P, P, S, P, S, P, ctypes.c_int]
        L.orc_expect_zstrings.argtypes = [P, S, P, P, S]
        L.orc_expect_zstrings.restype = ctypes.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _sizes(t) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(t, dtype=np.uint64).reshape(-1))


def _nqubits(state: np.ndarray) -> int:
    n = int(state.size).bit_length() - 1
    assert 1 << n == state.size, "state length must be a power of two"
    return n


def operation_valid(rows: int, cols: int, k: int) -> bool:
    return bool(lib().orc_operation_valid(rows, cols, k))


def new(bits) -> np.ndarray:
    bits = np.ascontiguousarray(np.asarray(bits, dtype=np.uint8).reshape(-1))
    out = np.zeros(1 << bits.size, dtype=np.complex128)
    lib().orc_new(_ptr(bits), bits.size, _ptr(out))
    return out


def from_int(n: int, k: int) -> np.ndarray:
    out = np.empty(1 << n, dtype=np.complex128)
    lib().orc_from_int(n, k, _ptr(out))
    return out


def try_from_qubits(qubits):
    """qubits: list of 1-D complex arrays.  Returns (code, state, a)."""
    lens = _sizes([np.asarray(q).size for q in qubits])
    flat = _c128(np.concatenate([np.asarray(q, dtype=np.complex128).reshape(-1) for q in qubits])
                 if qubits else np.zeros(0))
    n = len(qubits)
    out = np.zeros(1 << n, dtype=np.complex128)
    ea = ctypes.c_size_t(0)
    rc = lib().orc_try_from_qubits(_ptr(flat), _ptr(lens), n, _ptr(out), ctypes.byref(ea))
    return rc, (out if rc == OK else None), ea.value


def _apply(fn, state, matrix, targets):
    state = _c128(state).copy()
    m = _c128(matrix)
    rows, cols = (m.shape + (1, 1))[:2] if m.ndim == 2 else (m.size, 1)
    t = _sizes(targets)
    ea, eb = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = fn(_ptr(state), _nqubits(state), _ptr(m), rows, cols, _ptr(t), t.size,
            ctypes.byref(ea), ctypes.byref(eb))
    return rc, state, ea.value, eb.value


def try_apply_fast(state, matrix, targets):
    """Returns (code, new_state, a, b); new_state is a copy."""
    return _apply(lib().orc_try_apply_fast, state, matrix, targets)


def try_apply(state, matrix, targets):
    """Dense brute-force path; keep n <= 10."""
    return _apply(lib().orc_try_apply, state, matrix, targets)


def apply_fast(state, matrix, targets) -> np.ndarray:
    rc, out, a, b = try_apply_fast(state, matrix, targets)
    if rc != OK:
        raise ValueError(f"oracle apply failed: code {rc} ({a}, {b})")
    return out


def try_apply_all(state, matrix):
    state = _c128(state).copy()
    m = _c128(matrix)
    arity = (int(m.shape[0]).bit_length() - 1) if m.ndim == 2 else 0
    ea = ctypes.c_size_t(0)
    rc = lib().orc_try_apply_all(_ptr(state), _nqubits(state), _ptr(m), arity, ctypes.byref(ea))
    return rc, state, ea.value


def try_measure(state, target: int, x: float):
    """Returns (code, result_bool, new_state, a)."""
    state = _c128(state).copy()
    res = ctypes.c_int(0)
    ea = ctypes.c_size_t(0)
    rc = lib().orc_try_measure(_ptr(state), _nqubits(state), target, float(x), ctypes.byref(res),
                               ctypes.byref(ea))
    return rc, bool(res.value), state, ea.value


def measure_sequence(state, order, uniforms):
    """Measure the qubits in `order` one after another; returns (index, state)."""
    state = _c128(state).copy()
    o = _sizes(order)
    u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64).reshape(-1))
    assert u.size >= o.size
    idx = lib().orc_measure_sequence(_ptr(state), _nqubits(state), _ptr(o), o.size, _ptr(u))
    return int(idx), state


def apply_fast_omp(state: np.ndarray, scratch: np.ndarray, matrix, targets, threads: int) -> None:
    """In-place multi-core restatement of try_apply_fast (same arithmetic)."""
    assert state.dtype == np.complex128 and state.flags.c_contiguous
    assert scratch.dtype == np.complex128 and scratch.size >= state.size
    m = _c128(matrix)
    t = _sizes(targets)
    rc = lib().orc_apply_fast_omp(_ptr(state), _ptr(scratch), _nqubits(state), _ptr(m), t.size,
                                  _ptr(t), int(threads))
    if rc != 0:
        raise ValueError("arity too large for the multi-core oracle")


def expect_zstrings(state, masks, coeffs) -> float:
    state = _c128(state)
    mk = np.ascontiguousarray(np.asarray(masks, dtype=np.uint64))
    cf = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    return float(lib().orc_expect_zstrings(_ptr(state), _nqubits(state), _ptr(mk), _ptr(cf),
                                           mk.size))
