"""Gate values for the GPU register: quaru's ``Operation`` and its constructors.

Mirrors src/operation.rs of quaru 0.2.0.  An :class:`Operation` is a
``2^k x 2^k`` complex128 matrix plus ``targets``; matrix-index bit ``j``
addresses qubit ``targets[j]`` (src/register.rs:165-176), so ``cnot(c, t)``
stores ``targets == [t, c]`` (src/operation.rs:372-382).

Two additions the reference does not need:

* A *structured* form: ``Operation.controlled(controls, matrix, targets)``
  keeps the controls apart from the small matrix on the active targets.  It
  is what lets ``cnz`` / ``cnx`` on 32 qubits exist at all (the dense form
  would be 2^64 entries).  ``matrix()`` still materialises the dense
  equivalent (control = highest matrix bits, as ``to_controlled`` builds it,
  src/operation.rs:406-422) for registers small enough to hold it.
* ``ry`` / ``rz`` / ``rx`` rotation constructors, with the conventions of the
  OpenQASM ``U`` gate (src/openqasm.rs:145-163): ``u(t, 0, 0) == ry(t)`` and
  ``u(0, 0, l) == rz(l)``.
"""
from __future__ import annotations

import math
from typing import Callable, Iterable, List, Optional, Sequence

import numpy as np

_FRAC_1_SQRT_2 = 0.7071067811865476  # f64::consts::FRAC_1_SQRT_2


class Operation:
    """A quantum operation: a dense matrix and its targets (src/operation.rs:49-70).

    Use :meth:`new` (returns ``None`` on a bad shape, like ``Operation::new``)
    or one of the module's constructors.  ``controls`` is empty for a dense
    operation; for a structured one the full target list is
    ``targets + controls`` and ``matrix()`` is the dense equivalent.
    """

    __slots__ = ("_matrix", "_targets", "_controls", "_cvals", "_label")

    def __init__(self, matrix: np.ndarray, targets: Sequence[int],
                 controls: Sequence[int] = (), label: str = "",
                 control_values: Optional[Sequence[bool]] = None):
        self._matrix = np.ascontiguousarray(matrix, dtype=np.complex128)
        self._targets = [int(t) for t in targets]
        self._controls = [int(c) for c in controls]
        self._cvals = None if control_values is None else [bool(v) for v in control_values]
        if self._cvals is not None and len(self._cvals) != len(self._controls):
            raise ValueError("one control value per control")
        self._label = label

    # -- construction -------------------------------------------------------
    @staticmethod
    def new(matrix, targets: Sequence[int]) -> Optional["Operation"]:
        """``Operation::new`` (src/operation.rs:60-69): ``None`` unless the
        matrix is square with side ``2**len(targets)``."""
        m = np.asarray(matrix, dtype=np.complex128)
        d = 1 << len(targets)
        if m.ndim != 2 or m.shape[0] != d or m.shape[1] != d:
            return None
        return Operation(m, targets)

    @staticmethod
    def controlled(controls: Sequence[int], matrix, targets: Sequence[int],
                   label: str = "",
                   control_values: Optional[Sequence[bool]] = None) -> "Operation":
        """``matrix`` on ``targets``, applied only where qubit controls[i]
        equals control_values[i] (every control 1 when not given)."""
        m = np.asarray(matrix, dtype=np.complex128)
        d = 1 << len(targets)
        if m.shape != (d, d):
            raise ValueError("matrix side must be 2**len(targets)")
        return Operation(m, targets, controls, label, control_values)

    # -- QuantumOperation (src/operation.rs:34-38, 325-337) ------------------
    def matrix(self) -> np.ndarray:
        """The dense ``2^arity`` square matrix (a copy, as in the reference)."""
        if not self._controls:
            return self._matrix.copy()
        c = len(self._controls)
        k = len(self._targets)
        if k + c > 14:
            raise MemoryError(f"dense form of a {k + c}-qubit operation is too large")
        d_all = 1 << (k + c)
        full = np.eye(d_all, dtype=np.complex128)
        d = 1 << k
        cv = self._cvals if self._cvals is not None else [True] * c
        pattern = sum(1 << i for i, v in enumerate(cv) if v)
        base = pattern << k  # the control bits are the high matrix bits
        full[base:base + d, base:base + d] = self._matrix
        return full

    def targets(self) -> List[int]:
        return self._targets + self._controls

    def arity(self) -> int:
        return len(self._targets) + len(self._controls)

    # -- structured accessors used by the register ---------------------------
    @property
    def active_matrix(self) -> np.ndarray:
        return self._matrix

    @property
    def active_targets(self) -> List[int]:
        return list(self._targets)

    @property
    def controls(self) -> List[int]:
        return list(self._controls)

    @property
    def control_values(self) -> Optional[List[bool]]:
        return None if self._cvals is None else list(self._cvals)

    @property
    def is_structured(self) -> bool:
        return bool(self._controls)

    def __eq__(self, other) -> bool:  # derive(PartialEq)
        if not isinstance(other, Operation):
            return NotImplemented
        return (self.targets() == other.targets()
                and np.array_equal(self.matrix(), other.matrix()))

    def __repr__(self) -> str:
        name = self._label or "Operation"
        return f"{name}(targets={self._targets}, controls={self._controls})"


def _real(rows) -> np.ndarray:
    return np.asarray(rows, dtype=np.float64).astype(np.complex128)


def identity(target: int) -> Operation:
    """src/operation.rs:340-345"""
    return Operation(_real([[1.0, 0.0], [0.0, 1.0]]), [target], label="identity")


def hadamard(target: int) -> Operation:
    """src/operation.rs:350-355: FRAC_1_SQRT_2 * [[1, 1], [1, -1]]."""
    return Operation(_real([[_FRAC_1_SQRT_2, _FRAC_1_SQRT_2],
                            [_FRAC_1_SQRT_2, -_FRAC_1_SQRT_2]]), [target], label="hadamard")


def hadamard_transform(targets: Sequence[int]) -> Operation:
    """src/operation.rs:358-366: kron(H, acc) over the targets, so a later
    target is a higher matrix bit."""
    targets = list(targets)
    m = hadamard(targets[0]).matrix()
    for _t in targets[1:]:
        m = np.kron(hadamard(0).matrix(), m)
    return Operation(m, targets, label="hadamard_transform")


def cnot(control: int, target: int) -> Operation:
    """src/operation.rs:372-382: targets == [target, control]."""
    return Operation(_real([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
                     [target, control], label="cnot")


def to_quantum_gate(f: Callable[[int], int], targets: Sequence[int]) -> Operation:
    """src/operation.rs:386-401: column c of the matrix is e_{f(c)}."""
    d = 1 << len(targets)
    m = np.zeros((d, d), dtype=np.complex128)
    for c in range(d):
        m[f(c), c] = 1.0
    return Operation(m, list(targets), label="to_quantum_gate")


def to_controlled(op: Operation, control: int) -> Operation:
    """src/operation.rs:406-422: identity block first, ``op`` in the bottom
    right; the control becomes the new highest matrix bit."""
    old = 1 << op.arity()
    m = np.zeros((2 * old, 2 * old), dtype=np.complex128)
    m[:old, :old] = np.eye(old)
    m[old:, old:] = op.matrix()
    return Operation(m, op.targets() + [control], label="controlled")


def swap(target_1: int, target_2: int) -> Operation:
    """src/operation.rs:427-437"""
    return Operation(_real([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]),
                     [target_1, target_2], label="swap")


def phase(target: int) -> Operation:
    """src/operation.rs:443-451: diag(1, i)."""
    return Operation(np.array([[1, 0], [0, 1j]], dtype=np.complex128), [target], label="phase")


def not_(target: int) -> Operation:
    """src/operation.rs:458-463 (``not`` is a Python keyword)."""
    return Operation(_real([[0.0, 1.0], [1.0, 0.0]]), [target], label="not")


def pauli_y(target: int) -> Operation:
    """src/operation.rs:468-476"""
    return Operation(np.array([[0, -1j], [1j, 0]], dtype=np.complex128), [target], label="pauli_y")


def pauli_z(target: int) -> Operation:
    """src/operation.rs:481-486"""
    return Operation(_real([[1.0, 0.0], [0.0, -1.0]]), [target], label="pauli_z")


_DENSE_LIMIT = 12  # dense multi-controlled matrices up to 2^12 x 2^12


def cnx(controls: Sequence[int], target: int) -> Operation:
    """src/operation.rs:491-513: targets == [target, *controls]; identity with
    the last two rows swapped.  Built structured when the dense form would be
    large; both forms act identically."""
    controls = list(controls)
    k = 1 + len(controls)
    if k > _DENSE_LIMIT:
        return Operation.controlled(controls, _real([[0, 1], [1, 0]]), [target], label="cnx")
    d = 1 << k
    m = np.zeros((d, d), dtype=np.float64)
    for i in range(d - 2):
        m[i, i] = 1.0
    m[d - 1, d - 2] = 1.0
    m[d - 2, d - 1] = 1.0
    return Operation(m.astype(np.complex128), [target] + controls, label="cnx")


def cnz(controls: Sequence[int], target: int) -> Operation:
    """src/operation.rs:520-536: identity with -1 at the last diagonal entry."""
    controls = list(controls)
    k = 1 + len(controls)
    if k > _DENSE_LIMIT:
        return Operation.controlled(controls, _real([[1, 0], [0, -1]]), [target], label="cnz")
    d = 1 << k
    m = np.eye(d, dtype=np.complex128)
    m[d - 1, d - 1] = -1.0
    return Operation(m, [target] + controls, label="cnz")


def u(theta: float, phi: float, lam: float, target: int) -> Optional[Operation]:
    """OpenQASM U(theta, phi, lambda) (src/openqasm.rs:145-163)."""
    i = 1j
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    m = np.array([
        [np.exp(-i * (phi + lam) / 2.0) * c, -np.exp(-i * (phi - lam) / 2.0) * s],
        [np.exp(i * (phi - lam) / 2.0) * s, np.exp(i * (phi + lam) / 2.0) * c],
    ], dtype=np.complex128)
    return Operation.new(m, [target])


def ry(theta: float, target: int) -> Operation:
    """RY(theta) = exp(-i theta Y / 2) = u(theta, 0, 0)."""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return Operation(np.array([[c, -s], [s, c]], dtype=np.complex128), [target], label="ry")


def rz(theta: float, target: int) -> Operation:
    """RZ(theta) = exp(-i theta Z / 2) = u(0, 0, theta)."""
    return Operation(np.array([[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]],
                              dtype=np.complex128), [target], label="rz")


def rx(theta: float, target: int) -> Operation:
    """RX(theta) = exp(-i theta X / 2)."""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return Operation(np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128), [target],
                     label="rx")


def cz(a: int, b: int) -> Operation:
    """Controlled-Z on (a, b): ``cnz([a], b)``."""
    return cnz([a], b)


class QuantumCircuit:
    """Operation list + measurements + conditional operations
    (src/operation.rs:79-109, 271-322).  The reduce passes
    (src/operation.rs:111-269) are not reproduced: the register's own sweep
    planner fuses gates on the device instead."""

    def __init__(self):
        self.operations: List[Operation] = []
        self.measurement_targets: List[int] = []
        self.conditional_operations: List[tuple] = []

    def reset_register(self) -> "QuantumCircuit":
        self.operations.clear()
        self.measurement_targets.clear()
        self.conditional_operations.clear()
        return self

    def get_operations(self) -> List[Operation]:
        return self.operations

    def get_measurement_targets(self) -> List[int]:
        return self.measurement_targets

    def get_conditional_operations(self) -> List[tuple]:
        return self.conditional_operations

    def clear_operations(self) -> None:
        self.operations.clear()

    def add_operation(self, operation: Operation) -> None:
        """src/operation.rs:276-288"""
        if any(t in self.measurement_targets for t in operation.targets()):
            raise RuntimeError("Cannot add operation after measurement")
        if self.conditional_operations:
            raise RuntimeError("Cannot add operation after conditional operation (yet...)")
        self.operations.append(operation)

    def add_measurement(self, target: int) -> None:
        """src/operation.rs:292-297"""
        if target in self.measurement_targets:
            raise RuntimeError("Cannot add measurement after measurement")
        self.measurement_targets.append(target)

    def add_conditional_operation(self, operation: Operation,
                                  condition: Iterable[tuple]) -> None:
        """src/operation.rs:302-321"""
        condition = list(condition)
        if any(t in self.measurement_targets for t in operation.targets()):
            raise RuntimeError("Cannot add operation after measurement")
        if not all(c[0] in self.measurement_targets for c in condition):
            raise RuntimeError("Condition must be a measurement")
        self.conditional_operations.append((operation, condition))
