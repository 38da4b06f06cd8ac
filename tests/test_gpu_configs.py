"""The measured workloads at their full sizes on the GPU.

* 26 qubits, complex128 (1 GiB): the random-circuit construction of the
  12-qubit golden case.  The first two layers (plus merge chains) are
  compared per amplitude with the multi-core oracle (1e-12), at 26 qubits
  and moved onto qubits 4..29 of a 30-qubit register; all 20 layers,
  followed by their inverse, must return |0...0> within 1e-12 per amplitude
  (a mirror circuit; the state must first have left |0...0>).
* 30 qubits (16 GiB complex128 + 8 GiB complex64), 20 layers with 8
  Toffoli/CCZ per layer: the complex64 run against the complex128 run
  (infidelity <= 1e-6) and the complex128 mirror circuit back to |0...0>
  (1e-12 per amplitude, compared on the device over all 2^30 amplitudes).
* 32 qubits (64 GiB complex128, 32 GiB complex64): Grover with 12
  iterations, fused (5-qubit 32 x 32 Hadamard blocks) and unfused (one gate
  per qubit), against the closed form built on the device as its own
  register and compared over all 2^32 amplitudes (1e-12 for complex128,
  infidelity <= 1e-6 for complex64).
"""
import math
import os

import numpy as np
import pytest

import oracle
from quaru_amd import Register, circuits
from quaru_amd import operation as ops

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _zero(n: int, dtype: str = "c128") -> Register:
    return Register.new([False] * n, dtype)


_PREFIX26 = {}


def _prefix26():
    """The first two layers of the 26-qubit random circuit, then merge chains
    (a pattern qubit changed by a diagonal then a dense 2x2 between two ops
    of the same pattern), and the multi-core oracle's state after them."""
    if _PREFIX26:
        return _PREFIX26["ops"], _PREFIX26["ref"]
    from tests.test_gpu_register import merge_chain_ops
    n = 26
    w = circuits.random_circuit(n, 2, seed=2026, measure=False)
    opl = list(w.ops) + merge_chain_ops(np.random.default_rng(2626), n, 40)
    print(f"26q prefix: {len(opl)} gates on the oracle", flush=True)
    ref = np.zeros(1 << n, dtype=np.complex128)
    ref[0] = 1.0
    scratch = np.empty_like(ref)
    for op in opl:
        oracle.apply_fast_omp(ref, scratch, op.matrix(), op.targets(), _threads())
    _PREFIX26.update(ops=opl, ref=ref)
    return opl, ref


def test_random_26q_prefix_matches_oracle():
    n = 26
    opl, ref = _prefix26()
    r = _zero(n)
    for op in opl:
        r.apply(op)
    got = r.state
    assert float(np.max(np.abs(got - ref))) < TOL


def test_30q_high_qubits_match_oracle():
    """The 26-qubit prefix moved onto qubits 4..29 of a 30-qubit register
    (fusion on): the amplitudes whose low 4 index bits are zero must equal
    the oracle's 26-qubit state at 1e-12, and carry the whole norm (all
    others zero).  Qubits 26..29 thus get an amplitude-level check against
    the oracle, not only a mirror."""
    n, shift = 30, 4
    opl, ref = _prefix26()
    r = _zero(n)
    for op in opl:
        r.apply(ops.Operation(op.matrix(), [t + shift for t in op.targets()]))
    idx = (np.arange(1 << (n - shift), dtype=np.uint64) << np.uint64(shift))
    got = r.amplitudes(idx)
    assert float(np.max(np.abs(got - ref))) < TOL
    inside = float(np.sum(np.abs(got) ** 2))
    assert abs(inside - 1.0) < 1e-12 and abs(r.norm_sqr() - inside) < 1e-12
    assert abs(got[0]) < 0.5  # the state moved away from |0...0>


def test_random_26q_mirror_returns_to_zero():
    n = 26
    w = circuits.random_circuit(n, 20, seed=2026, measure=False)
    assert len(w.ops) == 26 * 20 + 25 * 20
    r = _zero(n)
    for op in w.ops:
        r.apply(op)
    assert abs(r.norm_sqr() - 1.0) < 1e-12
    assert abs(r.amplitudes([0])[0]) < 0.5  # the circuit moved the state
    for op in circuits.inverse(w.ops):
        r.apply(op)
    assert r.equals(_zero(n), TOL)


def _streaming_zoo(rng, n):
    """Per-gate kernels at a size where they stream (state above 256 MiB):
    one-qubit gates with partners inside a 128-byte line (qubits 0..2) and
    beyond it, two-qubit gates, 3..5-qubit blocks on low qubits (LDS-staged)
    and high ones (direct), complex and real, controlled gates, a diagonal
    and a sparse phase."""
    from tests.conftest import random_unitary

    def orth(d):
        q, r = np.linalg.qr(rng.standard_normal((d, d)))
        return (q * np.sign(np.diag(r))).astype(np.complex128)

    opl = [ops.Operation(random_unitary(rng, 2), [q]) for q in (0, 1, 2, 3, 4, 5, 12, n - 1)]
    opl += [ops.Operation(random_unitary(rng, 4), [1, n - 2]),
            ops.Operation(random_unitary(rng, 4), [7, 20])]
    for k in (3, 4, 5):
        opl += [ops.Operation(random_unitary(rng, 1 << k), list(range(k))),
                ops.Operation(random_unitary(rng, 1 << k), list(range(n - k, n))),
                ops.Operation(orth(1 << k), list(range(8, 8 + k)))]
    # one-qubit gates on qubits 0..2 under controls on bits >= 6 (the lane form)
    opl += [ops.to_controlled(ops.Operation(random_unitary(rng, 2), [1]), 20),
            ops.Operation.controlled([7, 18], random_unitary(rng, 2), [2],
                                     control_values=[True, False])]
    # complex128 one-qubit gates with a control on qubit 12 (the second pair
    # of a unit moves to qubit 11) and on 11..13 (no free qubit left for it:
    # one pair per unit)
    opl += [ops.to_controlled(ops.Operation(random_unitary(rng, 2), [20]), 12),
            ops.Operation.controlled([11, 12, 13], random_unitary(rng, 2), [21])]
    opl += [ops.cnot(0, n - 1), ops.cnx([3, 9], 17), ops.cnz([2, 5, 11], 23), ops.rz(0.7, 14),
            ops.Operation(np.diag(np.exp(1j * rng.uniform(0, 6, 8))), [4, 13, 22]),
            ops.to_controlled(ops.Operation(random_unitary(rng, 2), [n - 3]), 1)]
    return opl


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_streaming_per_gate_kernels_26q(dtype):
    """Fusion off at 26 qubits (1 GiB complex128, 512 MiB complex64): the
    per-gate kernels take their nontemporal form; every gate class against
    the multi-core oracle (1e-12 per amplitude, infidelity <= 1e-6 for
    complex64)."""
    n = 26
    rng = np.random.default_rng(2626)
    opl = _streaming_zoo(rng, n)
    r = Register.new([False] * n, dtype).set_fusion(False)
    r.apply_all(ops.hadamard(0))
    for op in opl:
        r.apply(op)
    ref = np.full(1 << n, 2.0 ** (-n / 2), dtype=np.complex128)
    scratch = np.empty_like(ref)
    for op in opl:
        oracle.apply_fast_omp(ref, scratch, op.matrix(), op.targets(), _threads())
    got = r.state
    if dtype == "c128":
        assert float(np.max(np.abs(got - ref))) < TOL
    else:
        inf = 1.0 - abs(np.vdot(ref, got)) ** 2 / (np.vdot(got, got).real * np.vdot(ref, ref).real)
        assert inf <= 1e-6


def test_random_30q_c64_vs_c128_and_mirror():
    n = 30
    w = circuits.random_circuit(n, 20, seed=2025, triples=8, measure=False)
    print(f"30q: {len(w.ops)} gates", flush=True)
    r128 = _zero(n)
    for op in w.ops:
        r128.apply(op)
    r64 = _zero(n, "c64")
    for op in w.ops:
        r64.apply(op)
    inf = r64.infidelity(r128)
    print(f"30q infidelity c64 vs c128: {inf:.3e}", flush=True)
    assert inf <= 1e-6
    del r64
    assert abs(r128.amplitudes([0])[0]) < 0.5  # the circuit moved the state
    for op in circuits.inverse(w.ops):
        r128.apply(op)
    assert r128.equals(_zero(n), TOL)


def test_33q_c64_mirror_64bit_indices():
    """33 qubits (64 GiB complex64): index arithmetic above 32 bits in the
    sweep and per-gate kernels; a 3-layer random circuit (with Toffoli/CCZ)
    and its inverse return |0...0>."""
    n = 33
    w = circuits.random_circuit(n, 3, seed=33, triples=4, measure=False)
    r = _zero(n, "c64")
    for op in w.ops:
        r.apply(op)
    assert abs(r.norm_sqr() - 1.0) < 1e-5
    probe = np.array([0, 1 << 32, (1 << 33) - 1, 5678901234], dtype=np.uint64)
    before = r.amplitudes(probe)
    for op in circuits.inverse(w.ops):
        r.apply(op)
    amps = r.amplitudes(probe)
    assert abs(amps[0] - 1.0) < 1e-5 and np.max(np.abs(amps[1:])) < 1e-6
    assert np.max(np.abs(before)) < 1.0  # the state really moved away from |0>
    # the whole state: |<0|psi>|^2 = 1 within complex64 rounding
    assert abs(abs(amps[0]) ** 2 / r.norm_sqr() - 1.0) < 1e-5


def _grover_closed_form_register(n: int, t: int, w: int, dtype: str) -> Register:
    """The state the closed form gives after t iterations, built on the
    device: ao * sum_x |x> (H on every qubit, then the scalar ao * 2^(n/2)
    as a 1-qubit multiple of I), then amplitude w scaled to aw."""
    aw, ao = circuits.grover_closed_form(n, t, w)
    e = Register.new([False] * n, dtype)
    e.apply_all(ops.hadamard(0))
    s = ao * 2.0 ** (n / 2)
    e.apply(ops.Operation(np.eye(2, dtype=np.complex128) * s, [0]))
    f = aw / ao
    m = np.diag([1.0, f]) if w & 1 else np.diag([f, 1.0])
    controls = list(range(1, n))
    e.apply(ops.Operation.controlled(controls, m.astype(np.complex128), [0],
                                     control_values=[bool((w >> q) & 1) for q in controls]))
    return e


def test_separable_blocks_split_into_sweeps():
    """Fusion on: separable dense blocks (the 5-qubit Hadamard blocks of the
    fused Grover workload, an 8-qubit hadamard_transform, a controlled
    product of two 2x2s) are applied as their one-qubit factors through the
    sweeps, with the oracle's result; an entangling block stays one gate."""
    from tests.conftest import random_unitary
    n, t = 14, 3
    w, opl = circuits.grover_operations(n, t, seed=9, fused=True)
    rng = np.random.default_rng(3)
    u, v = random_unitary(rng, 2), random_unitary(rng, 2)
    opl = opl + [ops.hadamard_transform(list(range(3, 11))),
                 ops.Operation.controlled([13], np.kron(u, v), [2, 9]),
                 ops.Operation(random_unitary(rng, 8), [1, 5, 12])]
    r = _zero(n)
    r.stats(reset=True)
    for op in opl:
        r.apply(op)
    got = r.state
    st = r.stats()
    ref = np.zeros(1 << n, dtype=np.complex128)
    ref[0] = 1.0
    for op in opl:
        ref = oracle.apply_fast(ref, op.matrix(), op.targets())
    assert float(np.max(np.abs(got - ref))) < TOL
    assert st["sweeps"] > 0 and st["gate_kernels"] <= 8, st  # the 3-qubit block and a few


def test_grover_closed_form_register_small():
    # the device-built comparison state is itself the closed form (n = 10)
    n, t = 10, 4
    w, opl = circuits.grover_operations(n, t, seed=9, fused=False)
    aw, ao = circuits.grover_closed_form(n, t, w)
    e = _grover_closed_form_register(n, t, w, "c128")
    want = np.full(1 << n, ao, dtype=np.complex128)
    want[w] = aw
    assert float(np.max(np.abs(e.state - want))) < 1e-14
    r = _zero(n)
    for op in opl:
        r.apply(op)
    assert r.equals(e, TOL)


@pytest.mark.parametrize("fused", [True, False])
def test_grover_32q_c128(fused):
    """fused: every Hadamard layer as 5-qubit 32 x 32 blocks, run by the block
    kernel (fusion off: with fusion on the engine would apply these separable
    blocks as their one-qubit factors through the sweeps, which the second
    test below checks at 14 qubits); unfused: one gate per qubit through the
    sweep engine."""
    n, t = 32, 12
    w, opl = circuits.grover_operations(n, t, seed=4, fused=fused)
    print(f"32q grover c128 fused={fused}: {len(opl)} ops, winner {w}", flush=True)
    r = _zero(n).set_fusion(not fused)
    r.stats(reset=True)
    for op in opl:
        r.apply(op)
    st = r.stats()
    assert (st["sweeps"] == 0) if fused else (st["sweeps"] > 0)
    aw, ao = circuits.grover_closed_form(n, t, w)
    rng = np.random.default_rng(3)
    idx = np.concatenate([[w, 0, (1 << n) - 1], rng.integers(0, 1 << n, 64)]).astype(np.uint64)
    amps = r.amplitudes(idx)
    want = np.where(idx == w, aw, ao)
    assert float(np.max(np.abs(amps - want))) < TOL
    e = _grover_closed_form_register(n, t, w, "c128")
    assert r.equals(e, TOL)


def test_grover_32q_c64():
    n, t = 32, 12
    w, opl = circuits.grover_operations(n, t, seed=4, fused=True)
    r = _zero(n, "c64").set_fusion(False)  # the 5-qubit blocks on the block kernel
    for op in opl:
        r.apply(op)
    e = _grover_closed_form_register(n, t, w, "c128")
    inf = r.infidelity(e)
    print(f"32q grover c64 infidelity vs closed form: {inf:.3e}", flush=True)
    assert inf <= 1e-6
    assert abs(r.norm_sqr() - 1.0) < 1e-5
