"""OpenQASM programs on the GPU register (src/openqasm.rs:52-142) against the
same programs executed by the oracle with the same measurement uniforms.

Bar: complex128 within 1e-12 per amplitude and identical classical
registers; complex64 within 1e-6 per amplitude and identical classical
registers on the programs whose outcome probabilities stay away from the
uniforms.
"""
import math

import numpy as np
import pytest

from quaru_amd import openqasm as qasm
from tests.test_openqasm import oracle_run

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-6}

GHZ = """
    OPENQASM 2.0;
    include "qelib1.inc";
    qreg q[5]; creg c[5];
    h q[0]; cx q[0], q[1]; cx q[1], q[2]; cx q[2], q[3]; cx q[3], q[4];
"""

TELEPORT = """
    include "qelib1.inc";
    qreg q[3]; creg m0[1]; creg m1[1];
    u3(0.8, 0.3, -1.1) q[0];
    h q[1]; cx q[1], q[2];
    cx q[0], q[1]; h q[0];
    measure q[0] -> m0[0]; measure q[1] -> m1[0];
    if (m1 == 1) x q[2];
    if (m0 == 1) z q[2];
"""


def qft_program(n):
    """QFT on n qubits from qelib1 gates (h, cu1, swap), then a measurement
    of every qubit, a reset and a second run of the prefix."""
    lines = ['include "qelib1.inc";', f"qreg q[{n}];", f"creg c[{n}];", "x q[0]; x q[2];"]
    for j in reversed(range(n)):
        lines.append(f"h q[{j}];")
        for k in reversed(range(j)):
            lines.append(f"cu1(pi/{2 ** (j - k)}) q[{k}], q[{j}];")
    for j in range(n // 2):
        lines.append(f"swap q[{j}], q[{n - 1 - j}];")
    lines += ["ccx q[0], q[1], q[2];", "measure q -> c;", "if (c == 5) x q[3];",
              "reset q;", "h q;", "rz(0.25) q[1];", "measure q[1] -> c[1];"]
    return "\n".join(lines)


def compare(prog, dtype, seed):
    res = qasm.run_program(prog, dtype, seed=seed)
    want_q, want_c = oracle_run(prog, seed)
    assert res.cregs == want_c
    for name, reg in res.qregs.items():
        d = float(np.max(np.abs(reg.state - want_q[name])))
        assert d < TOL[dtype], (name, d)
    return res


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_ghz_and_bell(dtype):
    res = compare(qasm.parse_openqasm_str(GHZ), dtype, 7)
    p = np.abs(res.qregs["q"].state) ** 2
    assert abs(p[0] - 0.5) < 1e-6 and abs(p[31] - 0.5) < 1e-6


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_teleportation_program(dtype):
    prog = qasm.parse_openqasm_str(TELEPORT)
    outcomes = set()
    for seed in range(8):
        res = compare(prog, dtype, seed)
        outcomes.add((res.cregs["m0"][0], res.cregs["m1"][0]))
    assert len(outcomes) >= 3


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_qft_measure_reset_program(dtype):
    compare(qasm.parse_openqasm_str(qft_program(10)), dtype, 11)


def test_cx_across_registers_is_a_run_error():
    prog = qasm.parse_openqasm_str('include "qelib1.inc"; qreg a[1]; qreg b[1]; cx a[0], b[0];')
    with pytest.raises(qasm.OpenQASMError) as e:
        qasm.run_program(prog)
    assert e.value.kind == "run"


def test_run_openqasm_file(tmp_path):
    path = tmp_path / "bell.qasm"
    path.write_text("""OPENQASM 2.0;
gate u2(phi, lambda) q { U(pi / 2, phi, lambda) q; }
gate h a { u2(0, pi) a; }
gate cx c, t { CX c, t; }
qreg q[3];
h q[0];
cx q[0], q[1];
""")
    res = qasm.run_openqasm(str(path))
    p = np.abs(res.qregs["q"].state) ** 2
    want = np.zeros(8)
    want[[0, 3]] = 0.5
    assert np.max(np.abs(p - want)) < 1e-7
    assert math.isclose(res.qregs["q"].norm_sqr(), 1.0, abs_tol=1e-12)
