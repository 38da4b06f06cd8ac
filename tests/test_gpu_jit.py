"""Sweep kernels specialised to an op-code sequence (compiled at run time with
hipRTC) against the generic sweep kernel and the CPU oracle: the same op
functions in the same order, so the states must agree to rounding (in
practice bit for bit), and both must match the oracle."""
import ctypes

import numpy as np
import pytest

from quaru_amd import Register, _lib, circuits
from quaru_amd.adjoint import Expectation, ParamCircuit
from tests.conftest import random_state
from tests.test_gpu_register import (TOL, gate_zoo, infid, maxdiff, merge_chain_ops,
                                     merge_stress_ops, oracle_run)

pytestmark = pytest.mark.gpu


def _stats():
    lib = _lib.load()
    c, f, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    assert lib.qs_sweep_jit_stats(ctypes.byref(c), ctypes.byref(f), ctypes.byref(n)) == 0
    return c.value, f.value, n.value


class _Mode:
    """Run a block with the given specialisation mode, then restore it."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = _lib.load().qs_sweep_jit_mode(self.mode)

    def __exit__(self, *exc):
        _lib.load().qs_sweep_jit_mode(self.prev)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_specialised_sweeps_match_generic_and_oracle(dtype):
    """Every op form the planner emits at 14 qubits (workgroup tiles with
    wave bits): the gate zoo, merge stress and merge chains (register, lane
    and wave targets, multiplexed forms, off-tile patterns, sign flips), then
    two layers of the random circuit."""
    n = 14
    rng = np.random.default_rng(91)
    opl = (gate_zoo(rng, n) + merge_stress_ops(rng, n, 40) + merge_chain_ops(rng, n, 20)
           + list(circuits.random_circuit(n, 2, seed=5, measure=False).ops))
    psi0 = random_state(rng, n)
    ref = oracle_run(psi0.copy(), opl)
    got = {}
    before = _stats()
    for mode in (0, 2):
        with _Mode(mode):
            r = Register.new([False] * n, dtype)
            r.state = psi0
            for op in opl:
                r.apply(op)
            got[mode] = r.state
            r.close()
    after = _stats()
    print(f"{dtype}: {after[0] - before[0]} kernels compiled, "
          f"{after[2] - before[2]} specialised launches", flush=True)
    assert after[1] == before[1], "a specialised kernel failed to compile"
    assert after[2] > before[2], "no specialised kernel ran"
    assert maxdiff(got[2], got[0]) < 1e-13
    if dtype == "c128":
        assert maxdiff(got[2], ref) < TOL
    else:
        assert infid(got[2], ref) < 1e-6


def test_specialised_batched_sweeps_adjoint():
    """The adjoint's forward pass (batched rotations folded per qubit, CZ sign
    flips) through specialised kernels: energies and gradients as with the
    generic kernel."""
    ans = circuits.hardware_efficient_ansatz(12, 3)
    masks, coeffs = circuits.zz_chain_observable(12)
    th = circuits.seeded_thetas(8, ans.nparams, seed=2)
    res = {}
    before = _stats()
    for mode in (0, 2):
        with _Mode(mode):
            ex = Expectation(ParamCircuit.from_ansatz(ans), masks, coeffs, checkpoint_every=1)
            e = ex.forward(th)
            res[mode] = (e, ex.backward())
    after = _stats()
    assert after[1] == before[1]
    assert after[2] > before[2]
    assert np.max(np.abs(res[2][0] - res[0][0])) < 1e-13
    assert np.max(np.abs(res[2][1] - res[0][1])) < 1e-12


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_specialised_sweeps_30q_benchmark_circuit(dtype):
    """The benchmark's own workload (30 qubits, 20 layers of the random
    circuit) through the specialised kernels it is timed on, against the
    generic kernel: one pass with the generic kernel queues the compiles
    (automatic mode), a second register then runs every sweep specialised.
    Same op functions in the same order, so the amplitudes must agree bit
    for bit (4096 sampled indices) and the overlap must be 1."""
    n = 30
    w = circuits.random_circuit(n, 20, seed=2025, triples=8, measure=False)
    idx = np.random.default_rng(7).integers(0, 1 << n, 4096, dtype=np.uint64)
    with _Mode(0):
        ra = Register.new([False] * n, dtype)
        for op in w.ops:
            ra.apply(op)
        ra.sync()
    with _Mode(1):
        rq = Register.new([False] * n, dtype)
        for op in w.ops:
            rq.apply(op)
        rq.sync()
        rq.close()
        assert _lib.load().qs_sweep_jit_wait() == 0
        before = _stats()
        rb = Register.new([False] * n, dtype)
        for op in w.ops:
            rb.apply(op)
        rb.sync()
        after = _stats()
    sweeps = rb.stats()["sweeps"]
    print(f"{dtype}: {after[2] - before[2]} specialised launches of {sweeps} sweeps", flush=True)
    assert after[1] == before[1], "a specialised kernel failed to compile"
    assert after[2] - before[2] == sweeps, "not every sweep ran specialised"
    a, b = ra.amplitudes(idx), rb.amplitudes(idx)
    assert np.array_equal(a, b), f"max diff {np.max(np.abs(a - b))}"
    norm = ra.norm_sqr()
    assert abs(ra.inner(rb) - norm) < 1e-12 * norm
    ra.close()
    rb.close()
