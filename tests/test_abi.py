"""CPU-side checks: the C-ABI library loads and exports every symbol that
include/quaru_hip.h declares (no compute calls), and the host-side logic
(gate constructors, workload generators) behaves."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from quaru_amd import _lib, circuits
from quaru_amd import operation as ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(qs_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_header_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 30
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    assert syms == set(_lib.SIGNATURES), syms ^ set(_lib.SIGNATURES)


def test_load_binds_signatures():
    lib = _lib.load()
    assert lib.qs_version().decode().startswith("quaru-hip")


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.LibraryMissing):
        _lib.load()


def all_ops():
    # src/operation.rs:547-563
    return [ops.identity(0), ops.hadamard(0), ops.cnot(0, 1), ops.swap(0, 1), ops.phase(0),
            ops.not_(0), ops.pauli_y(0), ops.pauli_z(0), ops.cnx([0], 1), ops.cnx([0, 1], 2),
            ops.cnx([0, 1, 2], 3), ops.cnx([0, 1, 2, 3], 4), ops.cnx([0, 1, 2, 3, 4], 5)]


def test_operation_shapes_and_unitarity():
    # src/operation.rs:565-584
    for op in all_ops():
        m = op.matrix()
        assert m.shape == (1 << op.arity(), 1 << op.arity())
        assert np.max(np.abs(m @ m.conj().T - np.eye(m.shape[0]))) < 1e-8


def test_toffoli2_equals_cnot():
    # src/operation.rs:586-594
    assert np.max(np.abs(ops.cnx([0], 1).matrix() - ops.cnot(0, 1).matrix())) < 1e-8
    assert ops.cnot(3, 5).targets() == [5, 3]
    assert ops.cnx([1, 2], 0).targets() == [0, 1, 2]


def test_operation_new_shape_check():
    assert ops.Operation.new(np.eye(4), [0, 1]) is not None
    assert ops.Operation.new(np.eye(2), [0, 1]) is None
    assert ops.Operation.new(np.eye(4), [0]) is None


def test_structured_equals_dense():
    s = ops.Operation.controlled([2, 3], [[0, 1], [1, 0]], [1])
    assert np.array_equal(s.matrix(), ops.cnx([2, 3], 1).matrix())
    big = ops.cnz(list(range(1, 32)), 0)
    assert big.is_structured and big.arity() == 32


def test_u_gate_conventions():
    # src/openqasm.rs:145-163: u(t, 0, 0) = RY(t), u(0, 0, l) = RZ(l)
    t = 0.7
    assert np.max(np.abs(ops.u(t, 0, 0, 0).matrix() - ops.ry(t, 0).matrix())) < 1e-15
    assert np.max(np.abs(ops.u(0, 0, t, 0).matrix() - ops.rz(t, 0).matrix())) < 1e-15
    h = ops.u(math.pi / 2, 0, math.pi, 0).matrix()  # the qasm example's h
    assert np.max(np.abs(np.abs(h) - np.abs(ops.hadamard(0).matrix()))) < 1e-12


def test_splitmix_known_values():
    # splitmix64 of seed 0: first outputs of the published sequence
    r = circuits.SplitMix64(0)
    assert r.next_u64() == 0xE220A8397B1DCDAF
    assert r.next_u64() == 0x6E789E6AA1B965F4
    u = circuits.SplitMix64(1).uniform()
    assert 0.0 <= u < 1.0


def test_haar_unitary_and_determinism():
    a = circuits.haar_2x2(circuits.SplitMix64(5))
    b = circuits.haar_2x2(circuits.SplitMix64(5))
    assert np.array_equal(a, b)
    assert np.max(np.abs(a @ a.conj().T - np.eye(2))) < 1e-14


def test_random_circuit_structure():
    w = circuits.random_circuit(6, 2, seed=1, triples=4)
    per_layer = 6 + 5 + 4
    assert len(w.ops) == 2 * per_layer == w.logical_gates
    tri = w.ops[11:15]
    assert [o._label for o in tri] == ["cnx", "cnz", "cnx", "cnz"]
    for o in tri:
        assert len(set(o.targets())) == 3
    assert w.changed[0] == 64 and w.changed[6] == 32 and w.changed[12] == 8
