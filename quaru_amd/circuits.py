"""Seeded synthetic workloads: the random-circuit, Grover and ansatz workloads.

Everything here is generated on the host in FP64 from one documented
counter-free generator, so the GPU engine and the CPU oracle receive the
very same gate matrices.

* :class:`SplitMix64` -- the splitmix64 sequence (Steele, Lea, Flood 2014):
  state += 0x9E3779B97F4A7C15, then two xor-shift-multiply rounds.
  ``uniform()`` = top 53 bits / 2^53 in [0, 1).
* :meth:`SplitMix64.normal` -- Box-Muller on two uniforms.
* :func:`haar_2x2` -- a Haar-random U(2): Z = (A + iB)/sqrt(2) with A, B
  i.i.d. N(0,1), then Gram-Schmidt on the columns of Z.  Gram-Schmidt gives
  the QR factor with a positive real diagonal R, which is exactly Mezzadri's
  phase-corrected Q (Q . diag(R_ii/|R_ii|)), the Haar measure on U(2).

Layer structure of the random circuits (the benchmark workloads): per layer a
fresh Haar unitary on every qubit 0..n-1, then ``cnot(q, q+1)`` for
q = 0..n-2, then (when asked for) ``triples`` three-qubit gates on seeded distinct
qubit triples (a, b, c): even gate index -> Toffoli ``cnx([a, b], c)``, odd
-> ``cnz([a, b], c)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import operation as ops

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = int(seed) & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal_pair(self) -> Tuple[float, float]:
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = 1.0 - u1  # (0, 1]: log is finite
        r = math.sqrt(-2.0 * math.log(u1))
        return r * math.cos(2.0 * math.pi * u2), r * math.sin(2.0 * math.pi * u2)

    def below(self, n: int) -> int:
        return min(int(self.uniform() * n), n - 1)


def haar_2x2(rng: SplitMix64) -> np.ndarray:
    g = [rng.normal_pair() for _ in range(4)]
    z = np.array([[complex(*g[0]), complex(*g[1])],
                  [complex(*g[2]), complex(*g[3])]], dtype=np.complex128) / math.sqrt(2.0)
    c0 = z[:, 0]
    q0 = c0 / np.sqrt(np.vdot(c0, c0).real)
    c1 = z[:, 1] - np.vdot(q0, z[:, 1]) * q0
    q1 = c1 / np.sqrt(np.vdot(c1, c1).real)
    return np.ascontiguousarray(np.stack([q0, q1], axis=1))


@dataclass
class Workload:
    """A gate list plus what the workloads attach to it."""
    n: int
    ops: List[ops.Operation]
    seed: int
    uniforms: List[float] = field(default_factory=list)
    logical_gates: int = 0
    # per-gate algorithmic amplitude count (changed set) for GB/s accounting
    changed: List[int] = field(default_factory=list)


def _changed(op: ops.Operation, n: int) -> int:
    """|changed set| of a gate: dense k-qubit 2^n, c controls 2^(n-c), a
    phase on the all-ones pattern of k qubits 2^(n-k)."""
    label = getattr(op, "_label", "")
    k = op.arity()
    if label in ("cnot", "cnx"):
        return 1 << (n - (k - 1))
    if label == "cnz":
        return 1 << (n - k)
    return 1 << n


def random_circuit(n: int, layers: int, seed: int, triples: int = 0,
                   measure: bool = True) -> Workload:
    rng = SplitMix64(seed)
    out: List[ops.Operation] = []
    for _layer in range(layers):
        for q in range(n):
            out.append(ops.Operation(haar_2x2(rng), [q], label="haar"))
        for q in range(n - 1):
            out.append(ops.cnot(q, q + 1))
        for g in range(triples):
            picked: List[int] = []
            while len(picked) < 3:
                q = rng.below(n)
                if q not in picked:
                    picked.append(q)
            a, b, c = picked
            out.append(ops.cnx([a, b], c) if g % 2 == 0 else ops.cnz([a, b], c))
    uniforms = [rng.uniform() for _ in range(n)] if measure else []
    w = Workload(n=n, ops=out, seed=seed, uniforms=uniforms, logical_gates=len(out))
    w.changed = [_changed(o, n) for o in out]
    return w


def inverse(opl: List[ops.Operation]) -> List[ops.Operation]:
    """The adjoint circuit: the conjugate transposes in reverse order, so that
    running opl and then inverse(opl) is the identity (a mirror circuit)."""
    out = []
    for op in reversed(opl):
        m = op.active_matrix.conj().T
        if op.is_structured:
            out.append(ops.Operation.controlled(op.controls, m, op.active_targets,
                                                control_values=op.control_values))
        else:
            out.append(ops.Operation(m, op.active_targets))
    return out


def grover_closed_form(n: int, t: int, w: int):
    """Amplitudes after H^n and t Grover iterations with the reference's signs:
    amp_w = sin((2t+1) th), others = cos((2t+1) th)/sqrt(N-1), sin th = N^-1/2."""
    N = float(1 << n)
    th = math.asin(1.0 / math.sqrt(N))
    aw = math.sin((2 * t + 1) * th)
    ao = math.cos((2 * t + 1) * th) / math.sqrt(N - 1.0)
    return aw, ao


def hadamard_layer(n: int, fused: bool, block: int = 5) -> List[ops.Operation]:
    """H on every qubit: n single-qubit gates, or (fused) dense Hadamard
    transforms on consecutive blocks of up to `block` qubits (32 x 32
    unitaries for block = 5)."""
    if not fused:
        return [ops.hadamard(q) for q in range(n)]
    return [ops.hadamard_transform(list(range(q, min(q + block, n))))
            for q in range(0, n, block)]


def phase_flip(n: int, index: int) -> ops.Operation:
    """-1 on basis state `index` only: diag on qubit 0 under controls on
    qubits 1..n-1 matching the other bits of `index`."""
    if n == 1:
        m = np.diag([-1.0, 1.0]) if index == 0 else np.diag([1.0, -1.0])
        return ops.Operation(m.astype(np.complex128), [0], label="flip")
    m = np.diag([1.0, -1.0]) if index & 1 else np.diag([-1.0, 1.0])
    controls = list(range(1, n))
    cvals = [bool((index >> q) & 1) for q in controls]
    return ops.Operation.controlled(controls, m, [0], label="flip", control_values=cvals)


def grover_operations(n: int, iterations: int, seed: int, fused: bool):
    """The Grover workload as Operations: H on all qubits, then `iterations`
    x [oracle: -1 on the seeded winner w; diffusion H^n . diag(1, -1, ..., -1)
    . H^n], with the signs of examples/grovers.rs:184-219.  diag(1, -1, ...,
    -1) is written as the global phase -1 (a 1-qubit -I) times a flip of
    basis state 0, so the per-amplitude result (global phase included) equals
    the dense matrices of the example.  Returns (w, operations)."""
    w = SplitMix64(seed).below(1 << n)
    minus_i = ops.Operation(-np.eye(2, dtype=np.complex128), [0], label="-I")
    out = hadamard_layer(n, fused)
    for _ in range(iterations):
        out.append(phase_flip(n, w))
        out += hadamard_layer(n, fused)
        out.append(phase_flip(n, 0))
        out.append(minus_i)
        out += hadamard_layer(n, fused)
    return w, out


@dataclass
class Ansatz:
    """Hardware-efficient ansatz of the differentiation config: per layer
    RY(theta) and RZ(theta) on every qubit, then a CZ ring (q, q+1 mod n)."""
    n: int
    layers: int
    # ops: ("ry"|"rz", qubit, param) or ("cz", a, b)
    ops: List[tuple]
    nparams: int
    layer_ends: List[int]


def hardware_efficient_ansatz(n: int, layers: int) -> Ansatz:
    out: List[tuple] = []
    p = 0
    ends = []
    for _ in range(layers):
        for q in range(n):
            out.append(("ry", q, p))
            p += 1
        for q in range(n):
            out.append(("rz", q, p))
            p += 1
        for q in range(n):
            a, b = q, (q + 1) % n
            if a != b:
                out.append(("cz", a, b))
        ends.append(len(out))
    return Ansatz(n=n, layers=layers, ops=out, nparams=p, layer_ends=ends)


def zz_chain_observable(n: int):
    """H = sum_i Z_i + 0.5 sum_{i=0}^{n-2} Z_i Z_{i+1} (open chain) as
    (masks, coeffs)."""
    masks = [1 << i for i in range(n)] + [(1 << i) | (1 << (i + 1)) for i in range(n - 1)]
    coeffs = [1.0] * n + [0.5] * (n - 1)
    return np.asarray(masks, dtype=np.uint64), np.asarray(coeffs, dtype=np.float64)


def seeded_thetas(batch: int, nparams: int, seed: int) -> np.ndarray:
    rng = SplitMix64(seed)
    return np.array([[2.0 * math.pi * rng.uniform() for _ in range(nparams)]
                     for _ in range(batch)], dtype=np.float64)
