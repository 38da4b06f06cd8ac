"""GPU parity: the HIP register against the CPU oracle.

Bar: complex128 results within 1e-12 per amplitude of the oracle (the
tolerance of the drop-in); complex64 within its own rounding
(infidelity against the FP64 oracle <= 1e-6).  Indices and outcomes are
compared exactly.  Every case runs with gate fusion on (the default sweep
engine) and off (one kernel per gate).
"""
import math

import numpy as np
import pytest

import oracle
from quaru_amd import OperationError, QuaruPanic, Register, circuits
from quaru_amd import operation as ops
from tests.conftest import random_state, random_unitary

pytestmark = pytest.mark.gpu

TOL = 1e-12


def oracle_run(state, opl):
    for op in opl:
        state = oracle.apply_fast(state, op.matrix(), op.targets())
    return state


def maxdiff(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def infid(a, b):
    ov = np.vdot(a, b)
    return 1.0 - abs(ov) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


# ---- construction --------------------------------------------------------------

@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_new_from_int_from_qubits(dtype):
    bits = [True, False, True, True, False, False, True]
    assert maxdiff(Register.new(bits, dtype).state, oracle.new(bits)) == 0
    assert maxdiff(Register.from_int(6, 37, dtype).state, oracle.from_int(6, 37)) == 0
    f = 1 / math.sqrt(2)
    qs = [[1, 0], [0, 1], [f, f], [f, -1j * f]]
    r = Register.from_qubits(qs, dtype)
    _, ref, _ = oracle.try_from_qubits(qs)
    assert maxdiff(r.state, ref) < (TOL if dtype == "c128" else 1e-7)
    with pytest.raises(OperationError) as e:
        Register.try_from_qubits([[1, 0], [1, 1]], dtype)
    assert e.value == OperationError.InvalidQubit(1)
    with pytest.raises(QuaruPanic):
        Register.from_qubits([[1, 0, 0]], dtype)


def test_zero_qubit_register():
    r = Register.new([])
    assert r.size() == 0 and maxdiff(r.state, [1]) == 0
    r.apply(ops.Operation.new(np.array([[1j]]), []))
    assert maxdiff(r.state, [1j]) == 0
    with pytest.raises(OperationError):
        r.try_measure(0)


# ---- validation (src/register.rs:141-153) ------------------------------------------

def test_validation_precedence_matches_oracle():
    r = Register.new([False] * 4)
    cases = [
        (np.eye(16), [9, 3, 1, 3]),
        (np.eye(16), [3, 1, 3, 1]),
        (np.eye(8), [1, 7, 5]),
        (np.eye(4), [2, 2]),
    ]
    for m, t in cases:
        rc, _, a, _ = oracle.try_apply_fast(oracle.new([False] * 4), m, t)
        with pytest.raises(OperationError) as e:
            r.try_apply_fast(ops.Operation(m, t))
        assert e.value == OperationError.InvalidTarget(a), (t, e.value)
    with pytest.raises(OperationError) as e:
        r.try_apply_fast(ops.Operation(np.eye(2), [0, 1]))
    assert e.value == OperationError.InvalidDimensions(2, 2)
    with pytest.raises(QuaruPanic):
        r.apply(ops.cnot(1, 4))
    with pytest.raises(OperationError) as e:
        r.try_apply_all(ops.cnot(0, 1))
    assert e.value == OperationError.InvalidArity(2)
    # src/register.rs:437-445
    r2 = Register.new([False, False])
    with pytest.raises(OperationError) as e:
        r2.try_apply(ops.cnot(1, 2))
    assert e.value == OperationError.InvalidTarget(2)
    r2.try_apply(ops.cnot(0, 1))


# ---- gate parity at small n, every gate class ------------------------------------

def gate_zoo(rng, n):
    """Gates exercising each classified kind and kernel."""
    zoo = []
    for q in rng.permutation(n)[:4]:
        zoo.append(ops.Operation(random_unitary(rng, 2), [int(q)]))
    t = [int(x) for x in rng.permutation(n)[:2]]
    zoo.append(ops.Operation(random_unitary(rng, 4), t))
    for k in (3, 4, 5, 6, 7):
        t = [int(x) for x in rng.permutation(n)[:k]]
        zoo.append(ops.Operation(random_unitary(rng, 1 << k), t))
    a, b, c, d = [int(x) for x in rng.permutation(n)[:4]]
    zoo += [ops.cnot(a, b), ops.cnx([a, b], c), ops.cnz([a, b], c), ops.cnz([a, b, c], d),
            ops.swap(a, c), ops.phase(b), ops.pauli_y(d), ops.pauli_z(a), ops.identity(c),
            ops.rz(0.3, b), ops.to_controlled(ops.Operation(random_unitary(rng, 4), [a, b]), c),
            ops.hadamard_transform([a, c, d])]
    diag = np.exp(1j * rng.uniform(0, 2 * np.pi, 8))
    zoo.append(ops.Operation(np.diag(diag), [b, d, a]))
    sparse = np.ones(16, dtype=np.complex128)
    sparse[5] = -1
    sparse[9] = 1j
    zoo.append(ops.Operation(np.diag(sparse), [c, a, d, b]))
    zoo.append(ops.Operation.new(np.array([[np.exp(0.4j)]]), []))
    return zoo


def random_orthogonal(rng, d):
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_block_gates_low_and_high_targets(dtype, real):
    """Dense 3..5-qubit blocks on low qubits (the LDS-staged kernel), on high
    qubits (the direct kernel) and under controls on low qubits; complex
    matrices and real ones (the 2-FMA form)."""
    rng = np.random.default_rng(11 if not real else 12)
    n = 14
    mat = (lambda d: random_orthogonal(rng, d)) if real else (lambda d: random_unitary(rng, d))
    opl = []
    for k in (3, 4, 5):
        opl.append(ops.Operation(mat(1 << k), list(range(k))))
        opl.append(ops.Operation(mat(1 << k), list(range(n - k, n))[::-1]))
        low = [int(x) for x in rng.permutation(4)[:2]]
        rest = [int(x) for x in rng.permutation(np.arange(4, n))[:k - 1]]
        opl.append(ops.Operation(mat(1 << k), [low[0]] + rest))
        opl.append(ops.to_controlled(ops.Operation(mat(1 << (k - 1)), rest[: k - 1]),
                                     low[1]))
    psi0 = random_state(rng, n)
    ref = oracle_run(psi0.copy(), opl)
    r = Register.new([False] * n, dtype)
    r.state = psi0
    for op in opl:
        r.apply(op)
    if dtype == "c128":
        assert maxdiff(r.state, ref) < TOL
    else:
        assert infid(r.state, ref) < 1e-6


@pytest.mark.parametrize("n", [12, 14])
def test_c64_complex_5q_blocks_matrix_cores(n):
    """Complex 32 x 32 blocks in complex64 with the 5 lowest index bits free
    run on the FP32 matrix cores (k_gatek5_mfma): targets in any order, with
    and without controls (above and between the targets), against the oracle
    (complex64 infidelity <= 1e-6 and per-amplitude error <= 2e-6).  The
    shor.rs QFT-add is a 5-qubit 32 x 32 known answer: QFT . add(a) . QFT^-1
    on |b> gives |a + b mod 32>."""
    rng = np.random.default_rng(40 + n)
    hi = list(range(5, n))
    # lowest active bit 5 (32 groups per wave) and >= 6 (64 groups per wave)
    opl = [ops.Operation(random_unitary(rng, 32), [9, 5, 7, 6, 8]),
           ops.Operation(random_unitary(rng, 32), [6, 7, 8, 9, 10])]
    for _ in range(3):
        t = [int(x) for x in rng.permutation(hi)[:5]]
        opl.append(ops.Operation(random_unitary(rng, 32), t))
    t = [int(x) for x in rng.permutation(hi)[:7]]
    opl.append(ops.to_controlled(ops.Operation(random_unitary(rng, 32), t[:5]), t[5]))
    opl.append(ops.Operation.controlled([t[5], t[6]], random_unitary(rng, 32), t[:5],
                                        control_values=[False, True]))
    psi0 = random_state(rng, n)
    ref = oracle_run(psi0.copy(), opl)
    r = Register.new([False] * n, "c64")
    r.state = psi0
    for op in opl:
        r.apply(op)
    got = r.state
    assert infid(got, ref) < 1e-6
    assert maxdiff(got, ref) < 2e-6
    # the QFT-add known answer (examples/shors.rs:466-503) on qubits 5..9
    qft = np.array([[np.exp(2j * np.pi * a * b / 32) for b in range(32)] for a in range(32)]) / np.sqrt(32)
    for a, b in ((3, 7), (17, 30), (31, 31)):
        add = np.diag([np.exp(2j * np.pi * a * k / 32) for k in range(32)])
        m = qft.conj().T @ add @ qft
        reg = Register.from_int(n, b << 5, "c64")
        reg.apply(ops.Operation(m, [5, 6, 7, 8, 9]))
        amp = reg.amplitudes([((a + b) % 32) << 5])[0]
        assert abs(abs(amp) - 1.0) < 1e-5, (a, b, amp)


@pytest.mark.parametrize("fusion", [True, False])
@pytest.mark.parametrize("n", [8, 12, 13])
def test_gate_zoo_parity_c128(n, fusion):
    rng = np.random.default_rng(100 + n)
    psi = random_state(rng, n)
    r = Register.new([False] * n).set_fusion(fusion)
    r.state = psi
    zoo = gate_zoo(rng, n)
    ref = psi
    for op in zoo:
        r.apply(op)
        ref = oracle.apply_fast(ref, op.matrix(), op.targets())
    assert maxdiff(r.state, ref) < TOL


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [12, 13, 14, 15])
def test_one_qubit_lane_form_chunk_orders(n, dtype):
    """Fusion off, one-qubit gates on qubits 0..5 with no control there: the
    lane kernel takes its four-stream chunk order from 128 chunks of 64
    amplitudes on (13 qubits uncontrolled, 14 with one control on a bit >= 6,
    15 with two) and its one-run order below; both against the oracle."""
    rng = np.random.default_rng(300 + n)
    psi = random_state(rng, n)
    r = Register.new([False] * n, dtype).set_fusion(False)
    r.state = psi
    opl = [ops.Operation(random_unitary(rng, 2), [q]) for q in range(6)]
    opl += [ops.to_controlled(ops.Operation(random_unitary(rng, 2), [q]), n - 1) for q in (0, 3, 5)]
    opl += [ops.Operation.controlled([6, n - 2], random_unitary(rng, 2), [q],
                                     control_values=[True, False]) for q in (1, 4)]
    ref = psi
    for op in opl:
        r.apply(op)
        ref = oracle.apply_fast(ref, op.matrix(), op.targets())
    got = r.state
    if dtype == "c128":
        assert maxdiff(got, ref) < TOL
    else:
        assert maxdiff(got, ref) < 1e-5


@pytest.mark.parametrize("n", [12, 14])
def test_gate_zoo_parity_c64(n):
    rng = np.random.default_rng(200 + n)
    psi = random_state(rng, n)
    r = Register.new([False] * n, "c64")
    r.state = psi
    ref = psi
    for op in gate_zoo(rng, n):
        r.apply(op)
        ref = oracle.apply_fast(ref, op.matrix(), op.targets())
    got = r.state
    assert infid(got, ref) < 1e-6
    assert maxdiff(got, ref) < 1e-5


# ---- the sweep engine on random circuits (the benchmark layer structure) ---------

@pytest.mark.parametrize("n,layers,triples", [(12, 4, 0), (14, 3, 4), (16, 2, 8), (18, 2, 8)])
def test_random_circuit_sweeps(n, layers, triples):
    w = circuits.random_circuit(n, layers, seed=n * 31 + layers, triples=triples, measure=False)
    ref = oracle_run(oracle.new([False] * n), w.ops)
    for fusion in (True, False):
        r = Register.new([False] * n).set_fusion(fusion)
        for op in w.ops:
            r.apply(op)
        st = r.stats()
        assert maxdiff(r.state, ref) < TOL, fusion
        if fusion:
            assert st["sweeps"] > 0
            assert st["gates"] == len(w.ops)


def merge_stress_ops(rng, n, count):
    """Runs of one-qubit gates mixed with the gates a sweep merges into them
    (CNOT / Toffoli targets, controlled 2x2s, controlled phases, CZ / CCZ),
    with targets and controls on every kind of tile position."""
    opl = []
    for _ in range(count):
        q = [int(x) for x in rng.permutation(n)[:4]]
        kind = rng.integers(0, 10)
        if kind < 3:
            opl.append(ops.Operation(random_unitary(rng, 2), [q[0]]))
        elif kind == 3:
            opl.append(ops.cnot(q[1], q[0]))
        elif kind == 4:
            opl.append(ops.cnx([q[1], q[2]], q[0]))
        elif kind == 5:
            opl.append(ops.cnz([q[1], q[2]], q[0]))
        elif kind == 6:
            opl.append(ops.cz(q[0], q[1]))
        elif kind == 7:
            opl.append(ops.to_controlled(ops.Operation(random_unitary(rng, 2), [q[0]]), q[1]))
        elif kind == 8:
            opl.append(ops.to_controlled(ops.phase(q[0]), q[1]))
        else:
            opl.append(ops.rz(float(rng.uniform(0, 6)), q[0]))
    return opl


def merge_chain_ops(rng, n, count):
    """Chains where a pattern qubit p changes between two ops of the same
    pattern on t: U(t); CX(p->t) (or a controlled 2x2, or CZ(p, t)); a
    diagonal 2x2 on p followed by a dense 2x2 on p, which merge into one plain
    op that is no longer diagonal; the same pattern op on t again.  The second
    pattern op must not fold back into the first across the op on p.  A few
    random one-qubit gates sit between chains."""
    opl = []
    while len(opl) < count:
        t, p, o = (int(x) for x in rng.permutation(n)[:3])
        kind = int(rng.integers(0, 3))

        def pat():
            if kind == 0:
                return ops.cnot(p, t)
            if kind == 1:
                return ops.cz(p, t)
            return ops.cnx([p, o], t)
        opl.append(ops.Operation(random_unitary(rng, 2), [t]))
        opl.append(pat())
        opl.append(ops.rz(float(rng.uniform(0, 6)), p))
        if rng.integers(0, 2):
            opl.append(ops.phase(p))
        opl.append(ops.Operation(random_unitary(rng, 2), [p]) if rng.integers(0, 2)
                   else ops.hadamard(p))
        opl.append(pat())
        if rng.integers(0, 2):
            opl.append(ops.Operation(random_unitary(rng, 2), [int(rng.integers(0, n))]))
    return opl


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [11, 14])
def test_sweep_merge_chains_match_oracle(n, dtype):
    """Merge chains across a pattern qubit (merge_chain_ops) give the
    oracle's result with fusion on."""
    rng = np.random.default_rng(901 + n)
    opl = merge_chain_ops(rng, n, 300)
    psi = random_state(rng, n)
    ref = oracle_run(psi.copy(), opl)
    r = Register.new([False] * n, dtype)
    r.state = psi
    r.stats(reset=True)
    for op in opl:
        r.apply(op)
    st = r.stats()
    assert st["merged"] > 20 and st["sweeps"] > 0
    if dtype == "c128":
        assert maxdiff(r.state, ref) < TOL
    else:
        assert infid(r.state, ref) < 1e-6


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [11, 14])
def test_sweep_merged_ops_match_oracle(n, dtype):
    """Merged multiplexed ops inside sweeps give the oracle's result."""
    rng = np.random.default_rng(77 + n)
    opl = merge_stress_ops(rng, n, 400)
    psi = random_state(rng, n)
    ref = oracle_run(psi.copy(), opl)
    r = Register.new([False] * n, dtype)
    r.state = psi
    r.stats(reset=True)
    for op in opl:
        r.apply(op)
    st = r.stats()
    assert st["merged"] > 40 and st["sweeps"] > 0
    if dtype == "c128":
        assert maxdiff(r.state, ref) < TOL
    else:
        assert infid(r.state, ref) < 1e-6


def test_golden_random_12q_on_gpu(golden):
    g = golden("c1_12q.npz")
    w = circuits.random_circuit(12, 20, seed=int(g["seed"]))
    for fusion in (True, False):
        r = Register.new([False] * 12).set_fusion(fusion)
        for op in w.ops:
            r.apply(op)
        assert maxdiff(r.state, g["amplitudes"]) < TOL
        idx = r.measure_sequence(range(12), w.uniforms)
        assert idx == int(g["index"])
        assert maxdiff(r.state, g["collapsed"]) < TOL


def test_golden_random_12q_c64(golden):
    g = golden("c1_12q.npz")
    w = circuits.random_circuit(12, 20, seed=int(g["seed"]))
    r = Register.new([False] * 12, "c64")
    for op in w.ops:
        r.apply(op)
    assert infid(r.state, g["amplitudes"]) < 1e-6


def test_golden_toffoli_mini(golden):
    g = golden("c3_mini.npz")
    w = circuits.random_circuit(10, 3, seed=7, triples=4, measure=False)
    r = Register.new([False] * 10)
    for op in w.ops:
        r.apply(op)
    assert maxdiff(r.state, g["amplitudes"]) < TOL


# ---- apply_all -------------------------------------------------------------------

@pytest.mark.parametrize("n", [2, 6, 13])
def test_apply_all(n):
    r = Register.new([False] * n)
    r.apply_all(ops.hadamard(0))
    assert maxdiff(r.state, np.full(1 << n, 2 ** (-n / 2))) < TOL
    rng = np.random.default_rng(n)
    u = random_unitary(rng, 2)
    psi = random_state(rng, n)
    r.state = psi
    r.apply_all(ops.Operation(u, [0]))
    ref = psi
    for q in range(n):
        ref = oracle.apply_fast(ref, u, [q])
    assert maxdiff(r.state, ref) < TOL


# ---- measurement -------------------------------------------------------------------

@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_measure_matches_oracle(dtype):
    rng = np.random.default_rng(11)
    n = 13
    psi = random_state(rng, n)
    for target in (0, 5, 12):
        for x in (0.1, 0.5, 0.93):
            r = Register.new([False] * n, dtype)
            r.state = psi
            res = r.measure(target, x)
            _, rr, ref, _ = oracle.try_measure(psi, target, x)
            assert res == rr
            tol = TOL if dtype == "c128" else 1e-6
            assert maxdiff(r.state, ref) < tol


@pytest.mark.parametrize("n", [13, 22, 25])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_measure_chain_with_pending_collapses(dtype, n):
    """Measures one after another (each collapse runs fused with the next
    measure's probability split) mixed with gates, readouts, a clone and a
    full readout, each of which must first see the collapse applied: the
    outcomes equal the oracle's on the same uniforms, the state after every
    step too (src/register.rs:291-342).  From 22 qubits on the split, the
    fused collapse and split and the norm take their block order (each
    workgroup a contiguous range), at 13 the grid-stride order; 25 qubits
    (512 MiB complex128) is a state beyond the last-level cache."""
    rng = np.random.default_rng(29)
    psi = random_state(rng, n)
    tol = TOL if dtype == "c128" else 1e-6
    r = Register.new([False] * n, dtype)
    r.state = psi
    ref = psi.copy()

    def meas(t, x):
        nonlocal ref
        got = r.measure(t, x)
        _, want, ref, _ = oracle.try_measure(ref, t, x)
        assert got == want, (t, x)

    meas(0, 0.31)
    meas(7, 0.62)            # fused with the collapse on qubit 0
    meas(12, 0.45)           # fused again
    assert maxdiff(r.amplitudes(np.arange(64)), ref[:64]) < tol   # readout settles
    meas(3, 0.77)
    meas(3, 0.05)            # the same target again: certain outcome
    h = ops.hadamard(5)
    r.apply(h)               # gate after a pending collapse: collapse first
    ref = oracle.apply_fast(ref, h.matrix(), h.targets())
    meas(5, 0.52)
    assert abs(r.norm_sqr() - float(np.vdot(ref, ref).real)) < (1e-9 if dtype == "c128" else 1e-6)
    meas(9, 0.2)
    c = r.clone()            # the clone sees the collapsed state
    assert maxdiff(c.state, ref) < tol
    c.close()
    meas(1, 0.66)
    u = [0.4, 0.9, 0.1]
    idx = r.measure_sequence([2, 4, 6], u)     # settles, then its own pass
    ridx, ref = oracle.measure_sequence(ref, [2, 4, 6], u)
    assert idx == ridx
    meas(11, 0.5)
    meas(n - 1, 0.4)
    meas(n - 2, 0.7)         # fused with the collapse on qubit n - 1
    assert abs(r.norm_sqr() - float(np.vdot(ref, ref).real)) < (1e-9 if dtype == "c128" else 1e-6)
    assert maxdiff(r.state, ref) < tol
    r.close()


def test_measure_basis_and_bell():
    rng = np.random.default_rng(3)
    for n in (4, 8):
        bits = [bool(b) for b in rng.integers(0, 2, n)]
        r = Register.new(bits, seed=1)
        assert [r.measure(q) for q in range(n)] == bits
    for i in range(6):
        j = (i + 2) % 6
        r = Register.new([False] * 6, seed=i)
        r.apply(ops.hadamard(i)).apply(ops.cnot(i, j))
        assert r.measure(i) == r.measure(j)


@pytest.mark.parametrize("n", [10, 14, 20, 25])
def test_measure_sequence_matches_oracle(n):
    rng = np.random.default_rng(n)
    psi = random_state(rng, n)
    for order in (list(range(n)), list(range(n))[::-1], [int(x) for x in rng.permutation(n)[:7]]):
        u = rng.random(len(order))
        r = Register.new([False] * n)
        r.state = psi
        idx = r.measure_sequence(order, u)
        ridx, ref = oracle.measure_sequence(psi, order, u)
        assert idx == ridx
        assert maxdiff(r.state, ref) < TOL


def test_measure_sequence_invalid_target_after_prefix():
    r = Register.new([True, False, True])
    with pytest.raises(OperationError) as e:
        r.measure_sequence([0, 1, 5, 2], [0.5] * 4)
    assert e.value == OperationError.InvalidTarget(5)
    assert maxdiff(r.state, oracle.new([True, False, True])) == 0


def test_sample_cdf():
    rng = np.random.default_rng(8)
    n = 18
    psi = random_state(rng, n)
    r = Register.new([False] * n)
    r.state = psi
    u = rng.random(64)
    got = r.sample(u)
    p = np.abs(psi) ** 2
    cdf = np.cumsum(p)
    exp = np.searchsorted(cdf, u * cdf[-1], side="right")
    assert np.array_equal(got, exp)


# ---- reductions and equality ---------------------------------------------------------

@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_readout_in_pieces(dtype):
    """At 24 qubits the readouts run through the register's work buffer in
    pieces of 2^22 amplitudes: set_state / state round trip (complex64
    narrows and widens per piece), probabilities, and 2^23 + 5 amplitudes by
    index, against numpy on the same values; a second call reuses the
    buffer."""
    n = 24
    rng = np.random.default_rng(24)
    psi = random_state(rng, n)
    r = Register.new([False] * n, dtype)
    r.state = psi
    want = psi if dtype == "c128" else psi.astype(np.complex64).astype(np.complex128)
    for _ in range(2):
        got = r.state
        assert np.array_equal(got, want)
        p = r.probabilities()
        assert np.max(np.abs(p - np.abs(want) ** 2)) < 1e-15
        idx = rng.integers(0, 1 << n, (1 << 23) + 5, dtype=np.uint64)
        assert np.array_equal(r.amplitudes(idx), want[idx])


def test_norm_inner_equal_probabilities():
    rng = np.random.default_rng(21)
    n = 15
    a, b = random_state(rng, n), random_state(rng, n)
    ra, rb = Register.new([False] * n), Register.new([False] * n)
    ra.state, rb.state = a, b
    assert abs(ra.norm_sqr() - 1.0) < 1e-13
    assert abs(ra.inner(rb) - np.vdot(a, b)) < 1e-13
    assert not (ra == rb)
    rc = ra.clone()
    assert ra == rc
    assert maxdiff(ra.probabilities(), np.abs(a) ** 2) < 1e-15
    ra64 = Register.new([False] * n, "c64")
    ra64.state = a
    assert abs(ra64.inner(ra) - 1.0) < 1e-6
    assert ra64.infidelity(ra) < 1e-10


# ---- reference property tests re-expressed on the GPU -------------------------------

def test_reference_properties():
    # tests/register_operation_tests.rs:44-50, 99-107, 143-181, 184-208
    for i in range(3):
        r = Register.new([False] * 3, seed=i)
        r.apply(ops.not_(i))
        assert r.measure(i)
        r = Register.new([False] * 3)
        expected = r.clone()
        h = ops.hadamard(i)
        assert r.apply(h).apply(h) == expected
    for op in (ops.hadamard(2), ops.not_(4), ops.identity(1), ops.cnot(1, 3), ops.swap(0, 5)):
        r = Register.new([False] * 6)
        expected = r.clone()
        r.apply(op).apply(op)
        assert r == expected
    for i in range(5):
        for j in range(5):
            if i == j:
                continue
            q = [False] * 6
            q[i] = True
            r = Register.new(q, seed=7)
            r.apply(ops.swap(i, j))
            assert [r.measure(k) for k in range(5)] == [k == j for k in range(5)]
    rng = np.random.default_rng(0)
    for n in range(2, 7):
        init = [bool(b) for b in rng.integers(0, 2, 6)]
        r = Register.new(init, seed=n)
        r.apply(ops.cnx(list(range(n - 1)), n - 1))
        ctl = all(r.measure(i) for i in range(n - 1))
        t = r.measure(n - 1)
        assert (t != init[n - 1]) if ctl else (t == init[n - 1])


def test_apply_circuit_teleportation():
    # examples/quantum_teleportation.rs:66-119 through apply_circuit
    rng = np.random.default_rng(4)
    for trial in range(6):
        th, ph = rng.uniform(0, math.pi), rng.uniform(0, 2 * math.pi)
        q0 = np.array([math.cos(th / 2), math.sin(th / 2) * np.exp(1j * ph)])
        r = Register.new([False] * 3, seed=trial)
        r.state = np.kron(np.kron([1, 0], [1, 0]), q0)
        c = ops.QuantumCircuit()
        for op in (ops.hadamard(2), ops.cnot(2, 1), ops.cnot(0, 1), ops.hadamard(0)):
            c.add_operation(op)
        c.add_measurement(1)
        c.add_measurement(0)
        c.add_conditional_operation(ops.not_(2), [(1, True)])
        c.add_conditional_operation(ops.pauli_z(2), [(0, True)])
        r.apply_circuit(c)
        s = r.state
        idx = np.arange(8)
        beta = s[(idx >> 2) & 1 == 1].sum()
        alpha = s[(idx >> 2) & 1 == 0].sum()
        p = abs(alpha) ** 2 + abs(beta) ** 2
        assert abs(alpha / math.sqrt(p) - q0[0]) < 1e-8
        assert abs(beta / math.sqrt(p) - q0[1]) < 1e-8


def test_grover_reference_ops_closed_form():
    # examples/grovers.rs with its dense k = n oracle and diffusion matrices
    n, t, w = 8, 5, 77
    N = 1 << n
    orc = np.eye(N)
    orc[w, w] = -1
    dif = -np.eye(N)
    dif[0, 0] = 1
    r = Register.new([False] * n)
    r.apply_all(ops.hadamard(0))
    for _ in range(t):
        r.apply(ops.Operation(orc, list(range(n))))
        r.apply_all(ops.hadamard(0))
        r.apply(ops.Operation(dif, list(range(n))))
        r.apply_all(ops.hadamard(0))
    aw, ao = circuits.grover_closed_form(n, t, w)
    s = r.state
    assert abs(s[w] - aw) < TOL and maxdiff(np.delete(s, w), ao) < TOL


def test_shor_slow_period_finding_register_ops():
    # examples/shors.rs:180-234 for N = 15, a = 7: 12 qubits, an 8-qubit
    # dense Hadamard transform, 8 controlled 5-qubit modmul gates and a dense
    # 256 x 256 QFT, compared with the oracle.
    number, a, nb = 15, 7, 4
    r = Register.new([False] * (3 * nb))
    opl = [ops.hadamard_transform(list(range(2 * nb))), ops.not_(2 * nb)]
    targets = list(range(2 * nb, 3 * nb))
    for i in range(2 * nb):
        apm = pow(a, 1 << i, number)
        g = ops.to_quantum_gate(lambda x, apm=apm: (x * apm) % number if x < number else x,
                                targets)
        opl.append(ops.to_controlled(g, i))
    m = 1 << (2 * nb)
    wq = np.exp(2j * math.pi / m)
    j = np.arange(m)
    opl.append(ops.Operation(wq ** np.outer(j, j) / math.sqrt(m), list(range(2 * nb))))
    for op in opl:
        r.apply(op)
    ref = oracle_run(oracle.new([False] * 12), opl)
    assert maxdiff(r.state, ref) < 1e-11


def _shor_u_gate(targets, number, a, i):
    """examples/shors.rs:79-92: x -> x * a^(2^i) mod N on the targets, built
    by to_quantum_gate from the same function (so arguments x >= N also map
    to x * a^(2^i) mod N, as in the reference)."""
    apm = pow(a, 1 << i, number)
    return ops.to_quantum_gate(lambda x: (x * apm) % number, list(targets))


def _shor_fast(backend, number, a, uniforms):
    """find_period_fast of examples/shors.rs:285-361 on a register-like
    backend: a control qubit reused 2n times (H, controlled U, the R phase
    from the exponentially weighted sum of earlier outcomes, H, measure,
    reset on 1).  Returns the measured integer."""
    n = math.ceil(math.log2(number + 1))
    backend.apply(ops.not_(1))
    res, sm = 0, 0.0
    targets = list(range(1, n + 1))
    for i in range(2 * n):
        backend.apply(ops.hadamard(0))
        backend.apply(ops.to_controlled(_shor_u_gate(targets, number, a, 2 * n - 1 - i), 0))
        phi = complex(math.cos(-2 * math.pi * sm), math.sin(-2 * math.pi * sm))
        backend.apply(ops.Operation(np.array([[1, 0], [0, phi]], dtype=np.complex128), [0]))
        backend.apply(ops.hadamard(0))
        m = backend.measure(0, float(uniforms[i]))
        sm /= 2.0
        if m:
            backend.apply(ops.not_(0))
            res |= 1 << i
            sm += 0.25
    return res


class _OracleRegister:
    """The oracle with the Register calls _shor_fast makes."""

    def __init__(self, n):
        self.state = oracle.new([False] * n)

    def apply(self, op):
        self.state = oracle.apply_fast(self.state, op.matrix(), op.targets())
        return self

    def measure(self, target, x):
        rc, r, self.state, _ = oracle.try_measure(self.state, target, x)
        assert rc == oracle.OK
        return bool(r)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("number,a", [(15, 7), (21, 2)])
def test_shor_fast_period_finding_replay(dtype, number, a):
    """Shor's fast period finder (examples/shors.rs:285-361, a measure-reset
    loop whose R-gate phase the host computes from earlier outcomes) on the
    GPU register and on the oracle with the same uniforms: the same outcome
    bits and the same final state.  Over the seeds the measured phase
    res / 2^(2n) lands near multiples of 1/r for the true period r."""
    from fractions import Fraction
    n = math.ceil(math.log2(number + 1))
    r_true = next(r for r in range(1, number) if pow(a, r, number) == 1)
    hits = 0
    for seed in range(6):
        rng = circuits.SplitMix64(1000 + seed)
        u = [rng.uniform() for _ in range(2 * n)]
        reg = Register.new([False] * (n + 1), dtype)
        got = _shor_fast(reg, number, a, u)
        ref = _OracleRegister(n + 1)
        want = _shor_fast(ref, number, a, u)
        assert got == want, (seed, got, want)
        assert maxdiff(reg.state, ref.state) < (TOL if dtype == "c128" else 2e-6)
        frac = Fraction(got, 1 << (2 * n)).limit_denominator(number - 1)
        hits += (r_true % frac.denominator == 0)
    assert hits >= 3


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_permutation_gates_gather(dtype):
    """Permutations with factors on more than 5 qubits run as a gather from a
    copy of the state (QS_K_PERM), controlled and uncontrolled, on low and
    high qubits; compared with the oracle."""
    import ctypes
    from quaru_amd import _lib
    from tests.test_plan_emulation import monomial_gate
    rng = np.random.default_rng(21)
    n = 14
    opl = []
    for k in (6, 8, 10):
        t = [int(x) for x in rng.permutation(n)[:k]]
        opl.append(monomial_gate(rng, t))
        opl.append(ops.Operation(random_unitary(rng, 2), [t[-1]]))
        opl.append(ops.to_controlled(monomial_gate(rng, t[1:]), t[0]))
    opl.append(ops.to_quantum_gate(lambda x: (x * 5) % 128 if x < 128 else x, list(range(6, 14))))
    psi = random_state(rng, n)
    ref = oracle_run(psi.copy(), opl)
    r = Register.new([False] * n, dtype)
    r.state = psi
    lib = _lib.load()
    assert lib.qs_timing_enable(r.handle, 1) == 0
    for op in opl:
        r.apply(op)
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_uint64 * 8)()
    assert lib.qs_timing_read(r.handle, ms, cnt, 1) == 0
    # every gate above except the controlled one on 5 active qubits
    assert cnt[_lib.QS_K_PERM] == 6
    if dtype == "c128":
        assert maxdiff(r.state, ref) < TOL
    else:
        assert infid(r.state, ref) < 1e-6


def circuit_with_conditions(rng, n):
    """A QuantumCircuit with operations, measurements of 3 qubits and
    conditional operations on the others (single, multiple, contradictory
    and empty conditions)."""
    c = ops.QuantumCircuit()
    q = [int(x) for x in rng.permutation(n)]
    meas, rest = q[:3], q[3:]
    for _ in range(12):
        t = [int(x) for x in rng.permutation(n)[:2]]
        c.add_operation(ops.Operation(random_unitary(rng, 4), t))
    for t in meas:
        c.add_measurement(t)
    c.add_conditional_operation(ops.not_(rest[0]), [(meas[0], True)])
    c.add_conditional_operation(ops.Operation(random_unitary(rng, 2), [rest[1]]),
                                [(meas[1], False), (meas[2], True)])
    c.add_conditional_operation(ops.pauli_z(rest[0]), [(meas[0], True), (meas[0], False)])
    c.add_conditional_operation(ops.Operation(random_unitary(rng, 4), rest[1:3]), [])
    c.add_conditional_operation(ops.cnot(rest[2], rest[0]), [(meas[2], False), (meas[2], False)])
    return c


def oracle_circuit(n, circuit, seed):
    """Register::apply_circuit (src/register.rs:345-366) on the oracle with
    the register's uniform stream."""
    rng = circuits.SplitMix64(seed)
    st = oracle.new([False] * n)
    for op in circuit.get_operations():
        st = oracle.apply_fast(st, op.matrix(), op.targets())
    measures = []
    for t in circuit.get_measurement_targets():
        rc, res, st, _ = oracle.try_measure(st, t, rng.uniform())
        measures.append((t, res))
    for op, cond in circuit.get_conditional_operations():
        if all((t, v) in measures for (t, v) in cond):
            st = oracle.apply_fast(st, op.matrix(), op.targets())
    return st, measures


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_apply_circuit_on_device_matches_oracle(dtype):
    """Measurements decided on the device and conditions applied as controls:
    the same outcomes as the oracle on the same uniforms, and the same state."""
    rng = np.random.default_rng(8)
    n = 9
    outcomes = set()
    for seed in range(10):
        c = circuit_with_conditions(rng, n)
        ref, want = oracle_circuit(n, c, seed)
        r = Register.new([False] * n, dtype, seed=seed)
        got = r.apply_circuit(c)
        assert got == want
        outcomes.add(tuple(v for _, v in got))
        if dtype == "c128":
            assert maxdiff(r.state, ref) < TOL
        else:
            assert maxdiff(r.state, ref) < 1e-5
    assert len(outcomes) >= 4


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_apply_circuit_condition_on_a_later_measured_target(dtype):
    """measure 0; X(1) if 0 measured 1; measure 1; Z(2) if 1 measured 0.  The
    second measurement is added after a conditional operation on its qubit
    (src/operation.rs:293-297 allows that), so the second condition reads the
    recorded outcome of qubit 1, not qubit 1 after the X: the oracle's
    apply_circuit on the same uniforms decides."""
    n = 4
    seen = set()
    for seed in range(12):
        c = ops.QuantumCircuit()
        c.add_operation(ops.hadamard(0))
        c.add_operation(ops.hadamard(1))
        c.add_operation(ops.hadamard(2))
        c.add_measurement(0)
        c.add_conditional_operation(ops.not_(1), [(0, True)])
        c.add_measurement(1)
        c.add_conditional_operation(ops.pauli_z(2), [(1, False)])
        ref, want = oracle_circuit(n, c, seed)
        r = Register.new([False] * n, dtype, seed=seed)
        got = r.apply_circuit(c)
        assert got == want
        seen.add(tuple(v for _, v in got))
        assert maxdiff(r.state, ref) < (TOL if dtype == "c128" else 1e-6)
    assert len(seen) == 4


def test_measure_deferred_equals_measure():
    import ctypes
    from quaru_amd import EngineError, _lib
    rng = np.random.default_rng(9)
    lib = _lib.load()
    for trial in range(6):
        psi = random_state(rng, 12)
        a = Register.new([False] * 12)
        b = Register.new([False] * 12)
        a.state = psi
        b.state = psi
        us = rng.uniform(0, 1, 12)
        res = [a.measure(t, float(us[t])) for t in range(12)]
        for t in range(12):
            assert lib.qs_measure_deferred(b.handle, t, float(us[t]), 11 - t) == 0
        out = (ctypes.c_ubyte * 12)()
        assert lib.qs_read_slots(b.handle, 0, 12, out) == 0
        assert [bool(out[11 - t]) for t in range(12)] == res
        assert maxdiff(a.state, b.state) < 1e-15
    r = Register.new([False] * 4)
    with pytest.raises(OperationError) as e:
        from quaru_amd.register import _check
        _check(lib.qs_measure_deferred(r.handle, 4, 0.5, 0))
    assert e.value == OperationError.InvalidTarget(4)
    with pytest.raises(EngineError):
        _check(lib.qs_measure_deferred(r.handle, 0, 0.5, _lib.QS_MEASURE_SLOTS))
    with pytest.raises(EngineError):
        _check(lib.qs_read_slots(r.handle, 60, 5, (ctypes.c_ubyte * 5)()))
