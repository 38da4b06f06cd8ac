"""Pin the CPU oracle (oracle/) before trusting it as the checker.

Every function of the oracle is checked against something that a plausible
mistake (dropped term, swapped index, transposed operand, wrong sign) would
fail: the golden fixtures (numpy tensor contraction, an independent
algorithm), the dense brute-force restatement of try_apply, closed forms,
and the reference's own property tests re-expressed
(tests/register_operation_tests.rs, examples/shors.rs, examples/grovers.rs,
examples/quantum_teleportation.rs, examples/superdense_coding.rs).
"""
import math

import numpy as np
import pytest

import oracle
from quaru_amd import circuits
from quaru_amd import operation as ops
from tests.conftest import random_state, random_unitary
from tests.golden.make_golden import apply_np


def run(state, opl):
    for op in opl:
        state = oracle.apply_fast(state, op.matrix(), op.targets())
    return state


# ---- construction (src/register.rs:41-102) -----------------------------------

@pytest.mark.parametrize("bits", [[], [True], [False, True, True], [True, False, True, True, False]])
def test_new_is_basis_state(bits):
    s = oracle.new(bits)
    k = sum(int(b) << i for i, b in enumerate(bits))
    e = np.zeros(1 << len(bits), dtype=np.complex128)
    e[k] = 1
    assert np.array_equal(s, e)


def test_from_int():
    s = oracle.from_int(5, 19)
    assert s[19] == 1 and np.count_nonzero(s) == 1


def test_from_qubits_kron_order():
    # tests/register_operation_tests.rs:25-40
    f = 1 / math.sqrt(2)
    rc, s, _ = oracle.try_from_qubits([[1, 0], [0, 1], [f, f]])
    assert rc == oracle.OK
    ref = oracle.apply_fast(oracle.new([False, True, False]), ops.hadamard(2).matrix(), [2])
    assert np.max(np.abs(s - ref)) < 1e-8


def test_from_qubits_invalid():
    f = 1 / math.sqrt(2)
    assert oracle.try_from_qubits([[1, 0], [f, 0.5]])[::2] == (oracle.INVALID_QUBIT, 1)
    assert oracle.try_from_qubits([[1, 0, 0], [1, 0]])[::2] == (oracle.INVALID_QUBIT, 0)
    # within 1e-6 of unit norm is accepted (src/register.rs:114)
    assert oracle.try_from_qubits([[1 + 4e-7, 0]])[0] == oracle.OK


# ---- try_apply_fast ------------------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_apply_fast_matches_dense_and_tensordot(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 8))
    k = int(rng.integers(1, min(n, 4) + 1))
    targets = [int(t) for t in rng.permutation(n)[:k]]
    m = random_unitary(rng, 1 << k)
    psi = random_state(rng, n)
    rc, fast, _, _ = oracle.try_apply_fast(psi, m, targets)
    assert rc == oracle.OK
    rc, dense, _, _ = oracle.try_apply(psi, m, targets)
    assert rc == oracle.OK
    npy = apply_np(psi, m, targets, n)
    assert np.max(np.abs(fast - dense)) < 1e-13
    assert np.max(np.abs(fast - npy)) < 1e-13


def test_transposed_matrix_is_caught():
    # A planted mistake (M^T instead of M) must fail the cross-checks.
    rng = np.random.default_rng(1)
    m = random_unitary(rng, 4)
    psi = random_state(rng, 5)
    good = apply_np(psi, m, [3, 1], 5)
    bad = oracle.apply_fast(psi, m.T, [3, 1])
    assert np.max(np.abs(good - bad)) > 1e-3
    swapped = oracle.apply_fast(psi, m, [1, 3])
    assert np.max(np.abs(good - swapped)) > 1e-3


def test_arity_zero_scales():
    psi = random_state(np.random.default_rng(0), 3)
    out = oracle.apply_fast(psi, np.array([[0.5j]]), [])
    assert np.max(np.abs(out - 0.5j * psi)) < 1e-15


def test_validation_precedence():
    psi = oracle.new([False] * 4)
    h = ops.hadamard(0).matrix()
    # 1. (rows, cols) == (k, k): a 2x2 matrix with two targets
    rc, _, a, b = oracle.try_apply_fast(psi, h, [0, 1])
    assert (rc, a, b) == (oracle.INVALID_DIMENSIONS, 2, 2)
    # 2. the smallest duplicate after sorting, before any range check
    rc, _, a, _ = oracle.try_apply_fast(psi, np.eye(16), [9, 3, 1, 3])
    assert (rc, a) == (oracle.INVALID_TARGET, 3)
    rc, _, a, _ = oracle.try_apply_fast(psi, np.eye(16), [3, 1, 3, 1])
    assert (rc, a) == (oracle.INVALID_TARGET, 1)
    # 3. first out-of-range target in targets order
    rc, _, a, _ = oracle.try_apply_fast(psi, np.eye(8), [1, 7, 5])
    assert (rc, a) == (oracle.INVALID_TARGET, 7)
    # src/register.rs:437-445: cnot(1, 2) on 2 qubits -> InvalidTarget(2)
    rc, _, a, _ = oracle.try_apply(oracle.new([False, False]), ops.cnot(1, 2).matrix(), [2, 1])
    assert (rc, a) == (oracle.INVALID_TARGET, 2)
    assert oracle.try_apply(oracle.new([False, False]), ops.cnot(1, 1).matrix(), [1, 1])[0] \
        == oracle.INVALID_TARGET
    assert oracle.operation_valid(4, 4, 2) and not oracle.operation_valid(2, 2, 2)


# ---- try_apply_all ----------------------------------------------------------------

@pytest.mark.parametrize("n", [2, 3, 5, 6])
def test_apply_all_equals_each_qubit(n):
    # tests/register_operation_tests.rs:211-219
    h = ops.hadamard(0).matrix()
    rc, s_all, _ = oracle.try_apply_all(oracle.new([False] * n), h)
    assert rc == oracle.OK
    s_each = run(oracle.new([False] * n), [ops.hadamard(q) for q in range(n)])
    assert np.max(np.abs(s_all - s_each)) < 1e-13
    assert np.max(np.abs(s_all - 2 ** (-n / 2))) < 1e-13  # closed form


def test_apply_all_random_unitary_and_arity():
    rng = np.random.default_rng(3)
    u = random_unitary(rng, 2)
    psi = random_state(rng, 5)
    rc, a, _ = oracle.try_apply_all(psi, u)
    b = psi
    for q in range(5):
        b = oracle.apply_fast(b, u, [q])
    assert rc == oracle.OK and np.max(np.abs(a - b)) < 1e-13
    rc, _, ea = oracle.try_apply_all(psi, ops.cnot(0, 1).matrix())
    assert (rc, ea) == (oracle.INVALID_ARITY, 2)


# ---- try_measure ------------------------------------------------------------------

@pytest.mark.parametrize("n", [4, 8])
def test_measure_basis_states_round_trip(n):
    # tests/register_operation_tests.rs:53-96
    rng = np.random.default_rng(n)
    for _ in range(8):
        bits = [bool(b) for b in rng.integers(0, 2, n)]
        s = oracle.new(bits)
        out = []
        for q in range(n):
            rc, r, s, _ = oracle.try_measure(s, q, rng.random())
            out.append(r)
        assert out == bits


def test_measure_rule_and_collapse():
    # amplitude 0.5 on |00>, sqrt(3)/2 on |01> -> prob_0 of qubit 0 = 0.25
    s = np.array([0.5, math.sqrt(3) / 2, 0, 0], dtype=np.complex128)
    rc, r, out, _ = oracle.try_measure(s, 0, 0.2)
    assert not r and np.allclose(out, [1, 0, 0, 0])
    rc, r, out, _ = oracle.try_measure(s, 0, 0.3)
    assert r and np.allclose(out, [0, 1, 0, 0])
    rc, r, out, _ = oracle.try_measure(s, 0, 0.25)  # strict '>'
    assert not r
    rc, _, _, a = oracle.try_measure(s, 2, 0.5)
    assert (rc, a) == (oracle.INVALID_TARGET, 2)


def test_bell_pairs_measure_equal():
    # tests/register_operation_tests.rs:110-140
    rng = np.random.default_rng(5)
    for i in range(6):
        for j in range(6):
            if i == j:
                continue
            s = run(oracle.new([False] * 6), [ops.hadamard(i), ops.cnot(i, j)])
            _, ri, s, _ = oracle.try_measure(s, i, rng.random())
            _, rj, s, _ = oracle.try_measure(s, j, rng.random())
            assert ri == rj


@pytest.mark.parametrize("n", range(2, 7))
def test_toffoli_cnx(n):
    # tests/register_operation_tests.rs:184-208
    rng = np.random.default_rng(n)
    for _ in range(6):
        init = [bool(b) for b in rng.integers(0, 2, 6)]
        s = run(oracle.new(init), [ops.cnx(list(range(n - 1)), n - 1)])
        ctl = all(init[:n - 1])
        k = sum(int(b) << i for i, b in enumerate(init))
        expect = k ^ (1 << (n - 1)) if ctl else k
        assert abs(s[expect] - 1) < 1e-12


# ---- known answers from the examples ------------------------------------------------

def qft(n):
    m = 1 << n
    w = np.exp(2j * math.pi / m)
    j = np.arange(m)
    return w ** np.outer(j, j) / math.sqrt(m)


def test_qft_add_identity():
    # examples/shors.rs:466-503 -- a ready-made dense 32x32 (k = 5) check
    n = 5
    q = qft(n)
    for a in range(0, 50, 7):
        add = np.eye(1)
        for i in range(n):
            em = np.array([[1, 0], [0, np.exp(1j * math.pi / 2 ** i * a)]])
            add = np.kron(add, em)
        for b in range(0, 50, 9):
            r1 = oracle.apply_fast(oracle.from_int(n, b % 32), q, list(range(n)))
            r1 = oracle.apply_fast(r1, add, list(range(n)))
            r2 = oracle.apply_fast(oracle.from_int(n, (a + b) % 32), q, list(range(n)))
            assert np.max(np.abs(r1 - r2)) < 1e-8


def test_u_gate_modmul_exhaustive():
    # examples/shors.rs:434-463: 4-qubit permutation gates, all cases
    n = 4
    for modulus in range(1, 16):
        for base in range(16):
            for i in range(0, 16, 5):
                apm = pow(base, 1 << i, modulus)
                g = ops.to_quantum_gate(lambda x: (x * apm) % modulus, list(range(n)))
                m = g.matrix()
                for inp in range(16):
                    out = oracle.apply_fast(oracle.from_int(n, inp), m, list(range(n)))
                    exp = (inp * apm) % modulus
                    assert abs(out[exp] - 1) < 1e-12


def grover_reference(n, t, w):
    """examples/grovers.rs: apply_all(H), then t x [dense oracle diag with -1 at
    w (k = n), diffuser = apply_all(H) . diag(1,-1..-1) . apply_all(H)]."""
    N = 1 << n
    h = ops.hadamard(0).matrix()
    s = oracle.try_apply_all(oracle.new([False] * n), h)[1]
    orc = np.eye(N)
    orc[w, w] = -1
    dif = -np.eye(N)
    dif[0, 0] = 1
    for _ in range(t):
        s = oracle.apply_fast(s, orc, list(range(n)))
        s = oracle.try_apply_all(s, h)[1]
        s = oracle.apply_fast(s, dif, list(range(n)))
        s = oracle.try_apply_all(s, h)[1]
    return s


@pytest.mark.parametrize("n,t", [(4, 1), (5, 3), (6, 4)])
def test_grover_closed_form(n, t):
    w = 5
    s = grover_reference(n, t, w)
    aw, ao = circuits.grover_closed_form(n, t, w)
    others = np.delete(s, w)
    assert abs(s[w] - aw) < 1e-12
    assert np.max(np.abs(others - ao)) < 1e-12


def test_teleportation_outcome_independent():
    # examples/quantum_teleportation.rs:47-119, 149-178
    rng = np.random.default_rng(9)
    for _ in range(10):
        th, ph = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
        q0 = np.array([math.cos(th / 2), math.sin(th / 2) * np.exp(1j * ph)])
        s = np.kron(np.kron([1, 0], [1, 0]), q0).astype(np.complex128)
        s = run(s, [ops.hadamard(2), ops.cnot(2, 1), ops.cnot(0, 1), ops.hadamard(0)])
        _, c1, s, _ = oracle.try_measure(s, 1, rng.random())
        if c1:
            s = oracle.apply_fast(s, ops.not_(2).matrix(), [2])
        _, c0, s, _ = oracle.try_measure(s, 0, rng.random())
        if c0:
            s = oracle.apply_fast(s, ops.pauli_z(2).matrix(), [2])
        idx = np.arange(8)
        beta = s[(idx >> 2) & 1 == 1].sum()
        alpha = s[(idx >> 2) & 1 == 0].sum()
        p = abs(alpha) ** 2 + abs(beta) ** 2
        assert abs(alpha / math.sqrt(p) - q0[0]) < 1e-8
        assert abs(beta / math.sqrt(p) - q0[1]) < 1e-8


def test_superdense_coding():
    # examples/superdense_coding.rs:85-96
    for a in (False, True):
        for b in (False, True):
            s = run(oracle.new([False, False]), [ops.hadamard(0), ops.cnot(0, 1)])
            enc = {(False, False): [ops.identity(0)], (False, True): [ops.not_(0)],
                   (True, False): [ops.pauli_z(0)],
                   (True, True): [ops.pauli_z(0), ops.not_(0)]}[(a, b)]
            s = run(s, enc + [ops.cnot(0, 1), ops.hadamard(0)])
            _, r0, s, _ = oracle.try_measure(s, 0, 0.5)
            _, r1, s, _ = oracle.try_measure(s, 1, 0.5)
            assert (r0, r1) == (a, b)


# ---- golden fixtures ------------------------------------------------------------------

def test_golden_random_12q(golden):
    g = golden("c1_12q.npz")
    w = circuits.random_circuit(12, 20, seed=int(g["seed"]))
    assert np.array_equal(np.asarray(w.uniforms), g["uniforms"])
    s = run(oracle.new([False] * 12), w.ops)
    assert np.max(np.abs(s - g["amplitudes"])) < 1e-12
    idx, coll = oracle.measure_sequence(s, range(12), w.uniforms)
    assert idx == int(g["index"])
    assert np.max(np.abs(coll - g["collapsed"])) < 1e-12


def test_golden_toffoli_mini(golden):
    g = golden("c3_mini.npz")
    w = circuits.random_circuit(10, 3, seed=7, triples=4, measure=False)
    s = run(oracle.new([False] * 10), w.ops)
    assert np.max(np.abs(s - g["amplitudes"])) < 1e-12


def test_omp_variant_matches_scalar():
    rng = np.random.default_rng(4)
    psi = random_state(rng, 10)
    for k in (1, 2, 3):
        m = random_unitary(rng, 1 << k)
        t = [int(x) for x in rng.permutation(10)[:k]]
        a = oracle.apply_fast(psi, m, t)
        b = psi.copy()
        oracle.apply_fast_omp(b, np.empty_like(b), m, t, 4)
        assert np.array_equal(a, b)


def test_expect_zstrings():
    rng = np.random.default_rng(2)
    psi = random_state(rng, 5)
    masks, coeffs = circuits.zz_chain_observable(5)
    x = np.arange(32)
    h = np.zeros(32)
    for mk, c in zip(masks, coeffs):
        h += c * (-1.0) ** np.array([bin(int(v) & int(mk)).count("1") for v in x])
    assert abs(oracle.expect_zstrings(psi, masks, coeffs) - np.sum(np.abs(psi) ** 2 * h)) < 1e-13
