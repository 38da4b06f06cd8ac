"""OpenQASM 2.0 front end (src/openqasm.rs:52-190), host side only.

The parser and the gate expansion run on the CPU; here the expanded basic
operations are executed by the oracle, so nothing touches the GPU.  What pins
them:

* the built-in qelib1 definitions expand to the textbook matrices (H, X, Y,
  Z, S, T, RX/RY/RZ, CZ, CY, SWAP, CH, Toffoli, CRZ, CU1, CU3) up to a global
  phase -- a wrong sign, a swapped parameter or a dropped gate in a
  decomposition fails here;
* closed forms: a Bell pair, GHZ by broadcast, teleportation with ``if``;
* ``creg_to_u32``'s bit order (src/openqasm.rs:175-189);
* the error cases of the reference's parser and runner.
"""
import math

import numpy as np
import pytest

import oracle
from quaru_amd import openqasm as qasm
from quaru_amd import operation as ops
from quaru_amd.circuits import SplitMix64

F = 1 / math.sqrt(2)


def f32(x):
    return float(np.float32(x))


def oracle_run(prog, seed=0):
    """Execute a program's basic operations on the oracle, drawing the
    measurement uniforms the way the GPU runner does (one SplitMix64 stream per
    qreg, seeded seed + rank of the name in sorted order)."""
    names = sorted(prog.qregs)
    states = {nm: oracle.new([False] * prog.qregs[nm]) for nm in names}
    rngs = {nm: SplitMix64(seed + k) for k, nm in enumerate(names)}
    cregs = {nm: [False] * s for nm, s in prog.cregs.items()}
    for cond, op in prog.basic_operations():
        if cond is not None and qasm.creg_to_int(cregs[cond[0]]) != cond[1]:
            continue
        kind = op[0]
        if kind == "U":
            (t, p, l), (reg, q) = op[1], op[2]
            g = ops.u(f32(t), f32(p), f32(l), q)
            states[reg] = oracle.apply_fast(states[reg], g.matrix(), g.targets())
        elif kind == "CX":
            (ra, a), (rb, b) = op[1], op[2]
            assert ra == rb
            g = ops.cnot(a, b)
            states[ra] = oracle.apply_fast(states[ra], g.matrix(), g.targets())
        elif kind == "measure":
            (qr, q), (cr, c) = op[1], op[2]
            rc, res, st, _ = oracle.try_measure(states[qr], q, rngs[qr].uniform())
            assert rc == oracle.OK
            states[qr] = st
            cregs[cr][c] = res
        elif kind == "resetq":
            states[op[1]] = oracle.new([False] * prog.qregs[op[1]])
        elif kind == "resetc":
            cregs[op[1]] = [False] * len(cregs[op[1]])
    return states, cregs


def unitary_of(prog_text, nq):
    """The unitary a gate-only program implements, column by column."""
    prog = qasm.parse_openqasm_str(prog_text)
    cols = []
    for b in range(1 << nq):
        st = oracle.from_int(nq, b)
        for cond, op in prog.basic_operations():
            assert cond is None
            if op[0] == "U":
                (t, p, l), (_, q) = op[1], op[2]
                g = ops.u(t, p, l, q)
            else:
                g = ops.cnot(op[1][1], op[2][1])
            st = oracle.apply_fast(st, g.matrix(), g.targets())
        cols.append(st)
    return np.array(cols).T


def equal_up_to_phase(a, b, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    k = np.argmax(np.abs(b))
    ph = a.flat[k] / b.flat[k]
    return abs(abs(ph) - 1) < tol and np.max(np.abs(a - ph * b)) < tol


def ctl(u):
    """Controlled-u with the control on qubit 0 (matrix bit 0) and u on qubit 1."""
    m = np.eye(4, dtype=complex)
    # index = b0 + 2 b1; control b0 = 1 -> indices 1 and 3
    m[np.ix_([1, 3], [1, 3])] = u
    return m


def rz(l):
    return np.diag([np.exp(-0.5j * l), np.exp(0.5j * l)])


def ry(t):
    return np.array([[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]])


def rx(t):
    return np.array([[math.cos(t / 2), -1j * math.sin(t / 2)],
                     [-1j * math.sin(t / 2), math.cos(t / 2)]])


H = np.array([[1, 1], [1, -1]]) * F
X = np.array([[0, 1], [1, 0]])
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1, -1])
S = np.diag([1, 1j])
T = np.diag([1, np.exp(0.25j * math.pi)])

ONE_QUBIT = [
    ("h", H), ("x", X), ("y", Y), ("z", Z), ("s", S), ("sdg", S.conj().T), ("t", T),
    ("tdg", T.conj().T), ("id", np.eye(2)), ("rx(0.7)", rx(0.7)), ("ry(-1.3)", ry(-1.3)),
    ("rz(2.1)", rz(2.1)), ("u1(0.4)", np.diag([1, np.exp(0.4j)])),
    ("u2(0.3,-0.9)", np.array([[1, -np.exp(-0.9j)], [np.exp(0.3j), np.exp(-0.6j)]]) * F),
]


@pytest.mark.parametrize("gate,want", ONE_QUBIT, ids=[g for g, _ in ONE_QUBIT])
def test_qelib1_one_qubit_gates(gate, want):
    u = unitary_of(f'include "qelib1.inc"; qreg q[1]; {gate} q[0];', 1)
    assert equal_up_to_phase(u, want)


SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])


def cu3(t, p, l):
    """RZ(p) RY(t) RZ(l): the target block of qelib1's cu3 (no extra phase)."""
    return rz(p) @ ry(t) @ rz(l)


TWO_QUBIT = [
    ("cx q[0],q[1];", ctl(X)), ("cz q[0],q[1];", ctl(Z)), ("cy q[0],q[1];", ctl(Y)),
    ("ch q[0],q[1];", ctl(H)), ("swap q[0],q[1];", SWAP),
    ("crz(0.9) q[0],q[1];", ctl(rz(0.9))), ("cu1(1.1) q[0],q[1];", ctl(np.diag([1, np.exp(1.1j)]))),
    ("cu3(0.5,1.2,-0.4) q[0],q[1];", ctl(cu3(0.5, 1.2, -0.4))),
]


@pytest.mark.parametrize("body,want", TWO_QUBIT, ids=[b.split()[0] for b, _ in TWO_QUBIT])
def test_qelib1_two_qubit_gates(body, want):
    u = unitary_of('include "qelib1.inc"; qreg q[2]; ' + body, 2)
    # the controlled gates carry no extra phase on the control-0 block
    assert np.max(np.abs(u - want)) < 1e-12 or equal_up_to_phase(u, want)


def test_qelib1_toffoli_is_exact():
    u = unitary_of('include "qelib1.inc"; qreg q[3]; ccx q[0],q[1],q[2];', 3)
    want = np.eye(8)
    want[[3, 7]] = want[[7, 3]]  # flip bit 2 where bits 0 and 1 are set
    assert equal_up_to_phase(u, want)
    # and it is the reference's cnx([0,1], 2)
    g = ops.cnx([0, 1], 2)
    ref = np.array([oracle.apply_fast(oracle.from_int(3, b), g.matrix(), g.targets())
                    for b in range(8)]).T
    assert equal_up_to_phase(u, ref)


def test_planted_sign_error_is_caught():
    # a cz built with the wrong phase gate is not a CZ
    bad = 'qreg q[2]; gate h a { U(pi/2,0,pi) a; } h q[1]; CX q[0],q[1]; U(0,0,pi/2) q[1]; h q[1];'
    assert not equal_up_to_phase(unitary_of(bad, 2), ctl(Z))


def test_expressions_and_parameters():
    prog = qasm.parse_openqasm_str("""
        OPENQASM 2.0;
        qreg q[1];
        gate g(a, b) x { U(a * 2 - b / 4, -pi / 2 + a ^ 2, sqrt(b) + cos(0) - ln(exp(1.5))) x; }
        g(0.5, 4.0) q[0];
        U(-(1+2)*3, 1e-1, .25) q[0];
    """)
    basic = prog.basic_operations()
    (t, p, l) = basic[0][1][1]
    assert t == pytest.approx(0.0) and p == pytest.approx(-math.pi / 2 + 0.25)
    assert l == pytest.approx(2.0 + 1.0 - 1.5)
    assert basic[1][1][1] == pytest.approx((-9.0, 0.1, 0.25))


def test_broadcast_measure_and_creg_order():
    prog = qasm.parse_openqasm_str("""
        include "qelib1.inc";
        qreg a[3]; qreg b[3]; creg c[3];
        x a; cx a, b; measure b -> c;
    """)
    basic = prog.basic_operations()
    assert sum(1 for _, o in basic if o[0] == "U") == 3
    cx = [o for _, o in basic if o[0] == "CX"]
    assert cx == [("CX", ("a", j), ("b", j)) for j in range(3)]
    assert [o for _, o in basic if o[0] == "measure"] == [
        ("measure", ("b", j), ("c", j)) for j in range(3)]
    assert qasm.creg_to_int([True, False, False]) == 1
    assert qasm.creg_to_int([False, True, True]) == 6
    with pytest.raises(qasm.OpenQASMError):
        qasm.creg_to_int([False] * 33)


def test_bell_pair_closed_form():
    prog = qasm.parse_openqasm_str("""
        OPENQASM 2.0;
        gate u2(phi, lambda) q { U(pi / 2, phi, lambda) q; }
        gate h a { u2(0, pi) a; }
        gate cx c, t { CX c, t; }
        qreg q[3];
        h q[0];
        cx q[0], q[1];
    """)
    st, _ = oracle_run(prog)
    p = np.abs(st["q"]) ** 2
    want = np.zeros(8)
    want[[0, 3]] = 0.5
    assert np.max(np.abs(p - want)) < 1e-7  # f32 angles: pi/2 is rounded


def test_teleportation_with_conditions():
    # teleport RY(0.8)|0> from q[0] to q[2]; corrections by classical ifs
    text = """
        include "qelib1.inc";
        qreg q[3]; creg m0[1]; creg m1[1];
        ry(0.8) q[0];
        h q[1]; cx q[1], q[2];
        cx q[0], q[1]; h q[0];
        measure q[0] -> m0[0]; measure q[1] -> m1[0];
        if (m1 == 1) x q[2];
        if (m0 == 1) z q[2];
    """
    prog = qasm.parse_openqasm_str(text)
    c, s = math.cos(f32(0.8) / 2), math.sin(f32(0.8) / 2)
    seen = set()
    for seed in range(12):
        st, cr = oracle_run(prog, seed)
        seen.add((cr["m0"][0], cr["m1"][0]))
        psi = st["q"].reshape(2, 2, 2)  # [q2, q1, q0]
        b0, b1 = int(cr["m0"][0]), int(cr["m1"][0])
        q2 = psi[:, b1, b0]
        assert equal_up_to_phase(q2, np.array([c, s]), tol=1e-6)
    assert len(seen) >= 3


def test_reset_and_errors():
    prog = qasm.parse_openqasm_str("""
        include "qelib1.inc";
        qreg q[2]; creg c[2];
        x q[0]; measure q -> c; reset q; reset c; h q[1];
    """)
    st, cr = oracle_run(prog)
    assert cr["c"] == [False, False]
    assert np.max(np.abs(np.abs(st["q"]) ** 2 - [0.5, 0, 0.5, 0])) < 1e-7
    bad = [
        "qreg q[2]; foo q[0];",                                   # unknown gate
        "qreg q[2]; U(0,0) q[0];",                                 # U arity
        "qreg q[2]; CX q[0], q[5];",                               # index out of range
        "qreg q[2]; CX q[1], q[1];",                               # one qubit twice
        "qreg q[2]; qreg q[3];",                                   # declared twice
        "qreg q[2]; creg c[3]; measure q -> c;",                   # sizes differ
        "qreg a[2]; qreg b[3]; CX a, b;",                          # broadcast sizes differ
        "OPENQASM 3.0; qreg q[1];",                                # version
        "qreg q[1]; U(0,0,0) q[0]",                                # missing ;
        "qreg q[1]; U(0,0,0) q[0]; $",                             # bad character
        'include "qelib1.inc"; qreg q[1]; rx(1,2) q[0];',          # parameter count
    ]
    for text in bad:
        with pytest.raises(qasm.OpenQASMError) as e:
            qasm.parse_openqasm_str(text).basic_operations()
        assert e.value.kind == "parse", text
    opaque = qasm.parse_openqasm_str("qreg q[1]; opaque magic a; magic q[0];")
    with pytest.raises(qasm.OpenQASMError) as e:
        opaque.basic_operations()
    assert e.value.kind == "run"


def test_include_relative_file(tmp_path):
    (tmp_path / "lib.inc").write_text("gate flip a { U(pi,0,pi) a; }\n")
    (tmp_path / "main.qasm").write_text('include "lib.inc"; qreg q[1]; flip q[0];\n')
    prog = qasm.parse_openqasm(str(tmp_path / "main.qasm"))
    st, _ = oracle_run(prog)
    assert abs(abs(st["q"][1]) - 1) < 1e-7
    with pytest.raises(qasm.OpenQASMError):
        qasm.parse_openqasm(str(tmp_path / "missing.qasm"))
