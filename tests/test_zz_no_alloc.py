"""qs_get_amplitudes_noalloc against qs_get_amplitudes, through the C ABI.

The no-allocation variant stages indices and amplitudes through one fixed
device array per GPU instead of hipMalloc.  Its amplitudes must equal the
allocating entry point's bit for bit, on two calls in a row with the second
one smaller (the array keeps the first call's contents, which must not leak
into the second), in both precisions; requests beyond its 65536-index limit
fail with QS_ERR_ARGUMENT.
"""
import ctypes

import numpy as np
import pytest

from quaru_amd import Register, _lib, circuits

pytestmark = pytest.mark.gpu


def _call(fn, reg, idx):
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.full(2 * idx.size, np.nan, dtype=np.float64)
    rc = fn(reg.handle, idx.ctypes.data_as(ctypes.c_void_p), idx.size,
            out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_amplitudes_equal_allocating_entry_point(dtype):
    lib = _lib.load()
    rng = np.random.default_rng(1234)
    for n, count in ((20, 65536), (14, 777)):  # the second call is smaller
        print(f"noalloc {dtype}: n={n} count={count}", flush=True)
        w = circuits.random_circuit(n, 3, seed=n, measure=False)
        r = Register.new([False] * n, dtype)
        for op in w.ops:
            r.apply(op)
        idx = rng.integers(0, 1 << n, count, dtype=np.uint64)
        idx[:3] = [0, (1 << n) - 1, idx[5]]  # ends of the state and a repeat
        rc_a, want = _call(lib.qs_get_amplitudes, r, idx)
        rc_b, got = _call(lib.qs_get_amplitudes_noalloc, r, idx)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(got, want)
        assert np.abs(want).max() > 0  # the state is not trivially zero


def test_noalloc_amplitudes_limits():
    lib = _lib.load()
    r = Register.new([False] * 6)
    rc, _ = _call(lib.qs_get_amplitudes_noalloc, r, np.zeros(65537, dtype=np.uint64))
    assert rc == _lib.QS_ERR_ARGUMENT
    rc, _ = _call(lib.qs_get_amplitudes_noalloc, r, np.array([64], dtype=np.uint64))
    assert rc == _lib.QS_ERR_ARGUMENT
    rc, out = _call(lib.qs_get_amplitudes_noalloc, r, np.array([0, 63], dtype=np.uint64))
    assert rc == 0 and np.array_equal(out, [1.0, 0.0, 0.0, 0.0])
