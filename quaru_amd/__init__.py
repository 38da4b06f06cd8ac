"""quaru_amd -- an MI355X (gfx950) state-vector engine for quaru's
gate-application path.

The GPU register keeps quaru's ``Register::new`` / ``apply`` / ``measure``
semantics (src/register.rs of quaru 0.2.0) over the C ABI in
``include/quaru_hip.h``; the state vector stays in HBM.  Gate values and
constructors mirror src/operation.rs.
"""
from . import operation
from .operation import (Operation, QuantumCircuit, cnot, cnx, cnz, cz, hadamard,
                        hadamard_transform, identity, not_, pauli_y, pauli_z, phase, rx, ry, rz,
                        swap, to_controlled, to_quantum_gate, u)
from .register import EngineError, OperationError, QuaruPanic, Register
from . import openqasm

__all__ = [
    "Register", "Operation", "QuantumCircuit", "OperationError", "QuaruPanic", "EngineError",
    "operation", "openqasm", "identity", "hadamard", "hadamard_transform", "cnot", "to_quantum_gate",
    "to_controlled", "swap", "phase", "not_", "pauli_y", "pauli_z", "cnx", "cnz", "cz", "u",
    "rx", "ry", "rz",
]
