"""The GPU register: quaru's ``Register`` with its state resident in HBM.

Mirrors src/register.rs of quaru 0.2.0 (``Register::new`` ... ``measure``)
over the C ABI of include/quaru_hip.h.  Semantics kept from the reference:

* qubit ``q`` is bit ``q`` of the basis index (src/register.rs:41-50);
* ``apply`` validates in the reference's order (dimensions, then the smallest
  duplicate target, then the first out-of-range target; src/register.rs:141-153)
  and returns ``self`` so calls chain (tests/register_operation_tests.rs:104);
* ``measure`` draws a uniform ``x`` and returns ``x > prob_0``
  (src/register.rs:317-321); the uniform comes from a seeded per-register
  generator instead of ``rand::thread_rng`` so runs are reproducible, or can
  be passed explicitly;
* ``try_*`` methods raise :class:`OperationError` where the reference returns
  ``Err``; the plain methods raise :class:`QuaruPanic` where it panics via
  ``expect`` (src/register.rs:75, :136, :261, :292);
* ``==`` is the reference's ``PartialEq``: every amplitude within 1e-8
  (src/register.rs:409-413).

The product path has no CPU fallback: constructing a register without the
built library or without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .circuits import SplitMix64
from .operation import Operation, QuantumCircuit


class OperationError(Exception):
    """``OperationError`` (src/register.rs:10-20).

    ``kind`` is one of ``InvalidTarget``, ``InvalidDimensions``,
    ``InvalidArity``, ``InvalidQubit``; ``values`` its payload.
    """

    _KINDS = {
        _lib.QS_INVALID_TARGET: ("InvalidTarget", 1),
        _lib.QS_INVALID_DIMENSIONS: ("InvalidDimensions", 2),
        _lib.QS_INVALID_ARITY: ("InvalidArity", 1),
        _lib.QS_INVALID_QUBIT: ("InvalidQubit", 1),
    }

    def __init__(self, kind: str, *values: int):
        super().__init__(f"{kind}({', '.join(str(v) for v in values)})")
        self.kind = kind
        self.values = tuple(int(v) for v in values)

    def __eq__(self, other):
        return (isinstance(other, OperationError) and self.kind == other.kind
                and self.values == other.values)

    def __hash__(self):
        return hash((self.kind, self.values))

    @staticmethod
    def InvalidTarget(t: int) -> "OperationError":
        return OperationError("InvalidTarget", t)

    @staticmethod
    def InvalidDimensions(r: int, c: int) -> "OperationError":
        return OperationError("InvalidDimensions", r, c)

    @staticmethod
    def InvalidArity(k: int) -> "OperationError":
        return OperationError("InvalidArity", k)

    @staticmethod
    def InvalidQubit(i: int) -> "OperationError":
        return OperationError("InvalidQubit", i)


class QuaruPanic(RuntimeError):
    """Raised where the reference panics (``expect`` on an ``Err``)."""


class EngineError(RuntimeError):
    """A HIP / device / argument failure of the engine itself."""

    def __init__(self, code: int, message: str):
        super().__init__(f"quaru-hip error {code}: {message}")
        self.code = code


def _check(rc: int) -> None:
    if rc == _lib.QS_OK:
        return
    a, b, msg = _lib.last_error()
    if rc in OperationError._KINDS:
        kind, nvals = OperationError._KINDS[rc]
        raise OperationError(kind, *((a, b)[:nvals]))
    raise EngineError(rc, msg)


def _dtype_code(dtype) -> int:
    if dtype in (_lib.QS_C128, "c128", "complex128", np.complex128):
        return _lib.QS_C128
    if dtype in (_lib.QS_C64, "c64", "complex64", np.complex64):
        return _lib.QS_C64
    raise ValueError(f"unknown dtype {dtype!r}")


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _sizes(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(list(v), dtype=np.uint64).reshape(-1))


def _vp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Register:
    """A quantum register of ``n`` qubits whose state lives on the GPU."""

    def __init__(self, _handle: int, seed: Optional[int] = None):
        # Use the class constructors (new, from_int, from_qubits).
        self._h = ctypes.c_void_p(_handle)
        if seed is None:
            seed = int.from_bytes(os.urandom(8), "little")
        self._rng = SplitMix64(seed)

    # -- construction (src/register.rs:41-102) -------------------------------
    @classmethod
    def new(cls, input_bits: Sequence[bool], dtype="c128", seed: Optional[int] = None):
        """``Register::new``: bit i of the input sets qubit i."""
        bits = np.ascontiguousarray(np.asarray([bool(b) for b in input_bits], dtype=np.uint8))
        h = ctypes.c_void_p()
        _check(_lib.load().qs_register_new(_vp(bits), bits.size, _dtype_code(dtype),
                                           ctypes.byref(h)))
        return cls(h.value, seed)

    @classmethod
    def from_int(cls, n: int, k: int, dtype="c128", seed: Optional[int] = None):
        """``Register::from_int``: the basis state |k> on n qubits."""
        h = ctypes.c_void_p()
        _check(_lib.load().qs_register_from_int(n, k, _dtype_code(dtype), ctypes.byref(h)))
        return cls(h.value, seed)

    @classmethod
    def try_from_qubits(cls, input_bits, dtype="c128", seed: Optional[int] = None):
        """``Register::try_from_qubits``: raises OperationError.InvalidQubit(i)."""
        arrs = [np.asarray(q, dtype=np.complex128).reshape(-1) for q in input_bits]
        lens = _sizes([a.size for a in arrs])
        flat = _c128(np.concatenate(arrs)) if arrs else np.zeros(1, dtype=np.complex128)
        h = ctypes.c_void_p()
        _check(_lib.load().qs_register_from_qubits(_vp(flat), _vp(lens), len(arrs),
                                                   _dtype_code(dtype), ctypes.byref(h)))
        return cls(h.value, seed)

    @classmethod
    def from_qubits(cls, input_bits, dtype="c128", seed: Optional[int] = None):
        try:
            return cls.try_from_qubits(input_bits, dtype, seed)
        except OperationError as e:
            raise QuaruPanic("Incorrect input qubits") from e

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if self._h is not None and self._h.value:
            _lib.load().qs_register_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clone(self) -> "Register":
        """``#[derive(Clone)]``: a deep copy on the device."""
        h = ctypes.c_void_p()
        _check(_lib.load().qs_register_clone(self._h, ctypes.byref(h)))
        r = Register(h.value, self._rng.next_u64())
        return r

    __copy__ = clone

    def __deepcopy__(self, memo):
        return self.clone()

    @property
    def handle(self):
        return self._h

    def size(self) -> int:
        """``Register::size``."""
        return int(_lib.load().qs_register_size(self._h))

    @property
    def dtype(self) -> str:
        return "c128" if _lib.load().qs_register_dtype(self._h) == _lib.QS_C128 else "c64"

    def seed(self, seed: int) -> "Register":
        """Reseed the generator of the measurement uniforms."""
        self._rng = SplitMix64(seed)
        return self

    # -- gates (src/register.rs:134-281) ---------------------------------------
    def try_apply_fast(self, op: Operation) -> "Register":
        L = _lib.load()
        t = _sizes(op.active_targets)
        m = _c128(op.active_matrix)
        if op.is_structured:
            c = _sizes(op.controls)
            cv = op.control_values
            cvp = None
            if cv is not None:
                cva = np.ascontiguousarray(np.asarray(cv, dtype=np.uint8))
                cvp = _vp(cva)
            _check(L.qs_apply_controlled(self._h, _vp(c), cvp, c.size, _vp(m), _vp(t), t.size))
        else:
            rows, cols = (m.shape + (0, 0))[:2] if m.ndim == 2 else (m.size, 1)
            _check(L.qs_apply(self._h, _vp(m), rows, cols, _vp(t), t.size))
        return self

    # try_apply (src/register.rs:193-250) computes the same map as
    # try_apply_fast; on the device both go through the same kernels.
    try_apply = try_apply_fast

    def apply(self, op: Operation) -> "Register":
        """``Register::apply``: panics (QuaruPanic) on an invalid operation."""
        try:
            return self.try_apply_fast(op)
        except OperationError as e:
            raise QuaruPanic("Could not apply operation") from e

    def apply_controlled(self, controls: Sequence[int], matrix, targets: Sequence[int],
                         control_values: Optional[Sequence[bool]] = None) -> "Register":
        """``matrix`` on ``targets`` where qubit controls[i] == control_values[i]
        (all 1 by default)."""
        t = _sizes(targets)
        c = _sizes(controls)
        m = _c128(matrix)
        cv = None
        if control_values is not None:
            cv = np.ascontiguousarray(np.asarray([bool(v) for v in control_values],
                                                 dtype=np.uint8))
        _check(_lib.load().qs_apply_controlled(self._h, _vp(c), None if cv is None else _vp(cv),
                                               c.size, _vp(m), _vp(t), t.size))
        return self

    def apply_diagonal(self, diag, targets: Sequence[int]) -> "Register":
        """The diagonal operation diag(d_0, ..., d_{2^k-1}) on ``targets``."""
        d = _c128(diag).reshape(-1)
        t = _sizes(targets)
        if d.size != 1 << t.size:
            raise OperationError.InvalidDimensions(d.size, d.size)
        _check(_lib.load().qs_apply_diagonal(self._h, _vp(d), _vp(t), t.size))
        return self

    def try_apply_all(self, op: Operation) -> "Register":
        """``Register::try_apply_all``: InvalidArity unless the op is unary."""
        m = _c128(op.matrix())
        rows, cols = m.shape
        _check(_lib.load().qs_apply_all(self._h, _vp(m), rows, cols, op.arity()))
        return self

    def apply_all(self, op: Operation) -> "Register":
        try:
            return self.try_apply_all(op)
        except OperationError as e:
            raise QuaruPanic("Could not apply operation") from e

    # -- measurement (src/register.rs:291-342) ---------------------------------
    def try_measure(self, target: int, uniform: Optional[float] = None) -> bool:
        x = self._rng.uniform() if uniform is None else float(uniform)
        res = ctypes.c_int(0)
        _check(_lib.load().qs_measure(self._h, int(target), x, ctypes.byref(res)))
        return bool(res.value)

    def measure(self, target: int, uniform: Optional[float] = None) -> bool:
        try:
            return self.try_measure(target, uniform)
        except OperationError as e:
            raise QuaruPanic("Could not measure bit") from e

    def measure_sequence(self, order: Sequence[int],
                         uniforms: Optional[Sequence[float]] = None) -> int:
        """Equivalent to ``measure(order[j], uniforms[j])`` for j in order;
        returns the basis index whose bit order[j] is outcome j."""
        o = _sizes(order)
        if uniforms is None:
            uniforms = [self._rng.uniform() for _ in range(o.size)]
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64).reshape(-1))
        if u.size < o.size:
            raise ValueError("need one uniform per measured qubit")
        idx = ctypes.c_uint64(0)
        _check(_lib.load().qs_measure_sequence(self._h, _vp(o), o.size, _vp(u),
                                               ctypes.byref(idx)))
        return int(idx.value)

    def measure_all(self, uniforms: Optional[Sequence[float]] = None) -> int:
        return self.measure_sequence(range(self.size()), uniforms)

    def sample(self, uniforms: Sequence[float]) -> np.ndarray:
        """Inverse-CDF samples (FP64 prefix sums) without collapsing."""
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64).reshape(-1))
        out = np.zeros(u.size, dtype=np.uint64)
        _check(_lib.load().qs_sample_cdf(self._h, _vp(u), u.size, _vp(out)))
        return out

    # -- circuits (src/register.rs:345-377) ------------------------------------
    def apply_circuit(self, circuit: QuantumCircuit) -> List[Tuple[int, bool]]:
        """``Register::apply_circuit`` (src/register.rs:345-366), kept on the
        device: the measurements are decided on the GPU (same uniforms, same
        rule as ``measure``), and each conditional operation is applied as an
        operation controlled on its condition's measured qubits.  After the
        collapse, "target measured v" holds exactly where qubit ``target`` is
        v, as long as no conditional operation acts on a measured qubit.
        ``add_conditional_operation`` only refuses targets measured *before*
        it (src/operation.rs:302-321); a later ``add_measurement`` may name a
        conditional operation's target, and then a condition on that qubit
        must read the recorded outcome, not the qubit: such circuits (and
        those with more measurements than device slots) are decided on the
        host.  One synchronisation at the end reads the outcomes."""
        for op in circuit.get_operations():
            self.apply(op)
        targets = list(circuit.get_measurement_targets())
        measured = set(int(t) for t in targets)
        if len(targets) > _lib.QS_MEASURE_SLOTS or any(
                measured.intersection(int(t) for t in op.targets())
                for op, _ in circuit.get_conditional_operations()):
            return self._apply_circuit_host(circuit)
        lib = _lib.load()
        for slot, t in enumerate(targets):
            _check(lib.qs_measure_deferred(self._h, int(t), self._rng.uniform(), slot))
        for op, cond in circuit.get_conditional_operations():
            want = {}
            for t, v in cond:
                want.setdefault(int(t), set()).add(bool(v))
            if any(len(v) > 1 for v in want.values()):
                continue  # (t, true) and (t, false): never both measured
            ctl = sorted(want)
            if not ctl:
                self.apply(op)
                continue
            self.apply_controlled(ctl, op.matrix(), op.targets(),
                                  [next(iter(want[t])) for t in ctl])
        out = np.zeros(max(len(targets), 1), dtype=np.uint8)
        _check(lib.qs_read_slots(self._h, 0, len(targets), _vp(out)))
        return [(int(t), bool(out[j])) for j, t in enumerate(targets)]

    def _apply_circuit_host(self, circuit: QuantumCircuit) -> List[Tuple[int, bool]]:
        # measure and decide on the host (src/register.rs:345-366 as written)
        measures: List[Tuple[int, bool]] = []
        if circuit.get_measurement_targets():
            measures = self.measure_circuit(circuit)
        for op, cond in circuit.get_conditional_operations():
            if all((t, v) in measures for (t, v) in cond):
                self.apply(op)
        return measures

    def measure_circuit(self, circuit: QuantumCircuit) -> List[Tuple[int, bool]]:
        return [(t, self.measure(t)) for t in circuit.get_measurement_targets()]

    # -- state access -----------------------------------------------------------
    @property
    def state(self) -> np.ndarray:
        """The ``pub state`` field: a host copy of the 2^n amplitudes."""
        out = np.empty(1 << self.size(), dtype=np.complex128)
        _check(_lib.load().qs_get_state(self._h, _vp(out)))
        return out

    @state.setter
    def state(self, value) -> None:
        v = _c128(value).reshape(-1)
        if v.size != 1 << self.size():
            raise ValueError("state length must be 2^n")
        _check(_lib.load().qs_set_state(self._h, _vp(v)))

    def get_state(self) -> np.ndarray:
        return self.state

    def amplitudes(self, indices) -> np.ndarray:
        """``state[i]`` for the given basis indices (complex128), without
        copying the whole state to the host."""
        idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64).reshape(-1))
        out = np.empty(idx.size, dtype=np.complex128)
        _check(_lib.load().qs_get_amplitudes(self._h, _vp(idx), idx.size, _vp(out)))
        return out

    def set_state(self, value) -> "Register":
        self.state = value
        return self

    def norm_sqr(self) -> float:
        out = ctypes.c_double(0.0)
        _check(_lib.load().qs_norm_sqr(self._h, ctypes.byref(out)))
        return out.value

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.size(), dtype=np.float64)
        _check(_lib.load().qs_probabilities(self._h, _vp(out)))
        return out

    def inner(self, other: "Register") -> complex:
        """<self|other> in FP64 (either storage precision)."""
        re, im = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _check(_lib.load().qs_inner(self._h, other._h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def infidelity(self, other: "Register") -> float:
        """1 - |<a|b>|^2 / (<a|a><b|b>), accumulated in FP64."""
        ov = self.inner(other)
        return 1.0 - (abs(ov) ** 2) / (self.norm_sqr() * other.norm_sqr())

    def equals(self, other: "Register", tol: float = 1e-8) -> bool:
        eq = ctypes.c_int(0)
        _check(_lib.load().qs_equal(self._h, other._h, float(tol), ctypes.byref(eq)))
        return bool(eq.value)

    def __eq__(self, other) -> bool:
        if not isinstance(other, Register):
            return NotImplemented
        return self.equals(other, 1e-8)

    __hash__ = None

    def sync(self) -> "Register":
        _check(_lib.load().qs_sync(self._h))
        return self

    def set_fusion(self, enabled: bool) -> "Register":
        _check(_lib.load().qs_set_fusion(self._h, 1 if enabled else 0))
        return self

    def stats(self, reset: bool = False) -> dict:
        s, k, g = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.load().qs_stats(self._h, ctypes.byref(s), ctypes.byref(k), ctypes.byref(g),
                                    1 if reset else 0))
        o, m = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.load().qs_stats_sweep_ops(self._h, ctypes.byref(o), ctypes.byref(m),
                                              1 if reset else 0))
        return {"sweeps": s.value, "gate_kernels": k.value, "gates": g.value,
                "sweep_ops": o.value, "merged": m.value}

    # -- printing (src/register.rs:379-401) --------------------------------------
    def print_probabilities(self) -> None:
        n = self.size()
        for i, p in enumerate(self.probabilities()):
            print(f"{i:0{n}b}: {p * 100.0}%")

    def print_nonzero_probabilities(self) -> None:
        n = self.size()
        for i, p in enumerate(self.probabilities()):
            if p > 0.0:
                print(f"{i:0{n}b}: {p * 100.0}%")

    def print_state(self) -> None:
        n = self.size()
        for i, s in enumerate(self.state):
            print(f"{i:0{n}b}: {s}")
