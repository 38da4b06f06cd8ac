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
