"""The sweep planner end to end on the CPU: the engine's classifier and
planner produce the exact sweep descriptors the GPU would run, and
tests/cpp/plan_emulator.cpp executes them on a host state the way k_sweep
decodes them (op code dispatch, slot tables, per-form pattern bits).  The
result must match the oracle to 1e-12 (descriptors of complex128 registers)
or to float rounding (complex64 descriptors, whose matrices are float).

This pins the planner's gate ordering, the op merging, the slot encoding and
the descriptor layout without a GPU; the kernels' arithmetic is pinned by the
GPU tests.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from quaru_amd import circuits
from quaru_amd import operation as ops
from tests.conftest import random_state, random_unitary
from tests.test_gpu_register import gate_zoo, merge_chain_ops, merge_stress_ops

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "quaru_amd", "csrc")
LIBDIR = os.path.join(ROOT, "quaru_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "plan_emulator.cpp")
HIPCC = "/opt/rocm/bin/hipcc"

pytestmark = pytest.mark.skipif(
    not (os.path.exists(HIPCC) and os.path.exists(os.path.join(LIBDIR, "libquaru_hip.so"))),
    reason="needs hipcc and the built library")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu") / "libplan_emulator.so")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O1", "-fPIC",
                    "-shared", "-I", CSRC, SRC, "-L", LIBDIR, "-lquaru_hip",
                    f"-Wl,-rpath,{LIBDIR}", "-o", out],
                   check=True, capture_output=True, timeout=600)
    lib = ctypes.CDLL(out)
    lib.emu_run.restype = ctypes.c_int
    lib.emu_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int] + \
        [ctypes.c_void_p] * 7 + [ctypes.c_char_p, ctypes.c_int]

    def run(n, dtype, psi, opl):
        """psi None: plan only.  Structured (controlled) operations pass their
        active matrix and controls, as Register.apply does."""
        st = None if psi is None else np.ascontiguousarray(psi, dtype=np.complex128).copy()
        mats, tg, ks, cms, cvs = [], [], [], [], []
        for op in opl:
            cm = cv = 0
            if op.is_structured:
                m, t = op.active_matrix, op.active_targets
                vals = op.control_values or [True] * len(op.controls)
                for c, v in zip(op.controls, vals):
                    cm |= 1 << c
                    cv |= (1 << c) if v else 0
            else:
                m, t = op.matrix(), op.targets()
            mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
            tg.append(np.asarray(t, dtype=np.uint64))
            ks.append(len(t))
            cms.append(cm)
            cvs.append(cv)
        mats = np.ascontiguousarray(np.concatenate(mats))
        tg = np.ascontiguousarray(np.concatenate(tg), dtype=np.uint64)
        ks = np.ascontiguousarray(ks, dtype=np.int32)
        cms = np.ascontiguousarray(cms, dtype=np.uint64)
        cvs = np.ascontiguousarray(cvs, dtype=np.uint64)
        stats = np.zeros(4, dtype=np.uint64)
        hist = np.zeros(128, dtype=np.uint64)
        err = ctypes.create_string_buffer(512)
        rc = lib.emu_run(n, 1 if dtype == "c64" else 0, None if st is None else st.ctypes.data,
                         len(opl), mats.ctypes.data, tg.ctypes.data, ks.ctypes.data,
                         cms.ctypes.data, cvs.ctypes.data, stats.ctypes.data, hist.ctypes.data,
                         err, 512)
        assert rc == 0, err.value.decode()
        out = dict(zip(["sweeps", "singles", "sweep_ops", "merged"], map(int, stats)))
        out["hist"] = hist
        return st, out
    return run


def oracle_run(psi, opl):
    for op in opl:
        psi = oracle.apply_fast(psi, op.matrix(), op.targets())
    return psi


TOL = {"c128": 1e-12, "c64": 2e-6}


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [11, 14, 16])
def test_merge_stress_plans(emu, n, dtype):
    rng = np.random.default_rng(77 + n)
    opl = merge_stress_ops(rng, n, 400)
    psi = random_state(rng, n)
    got, st = emu(n, dtype, psi, opl)
    ref = oracle_run(psi.copy(), opl)
    assert st["sweeps"] > 0 and st["merged"] > 40
    assert np.max(np.abs(got - ref)) < TOL[dtype]


def test_batched_rotations_plan(tmp_path):
    """The adjoint's fused forward pass: a batch of 8 states stacked in index
    bits 13..15 of a 16-qubit register, layers of per-state RY/RZ rotations
    (batched 2x2s: the kernel picks state b's matrix by the tile's batch
    bits) and a CZ ring.  The planner must keep the batch bits off the tile,
    fold each RZ(q) into the RY(q) before it, and every state must equal its
    own circuit on the oracle."""
    out = str(tmp_path / "libplan_emulator.so")
    subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-std=c++17", "-O1", "-fPIC",
                    "-shared", "-I", CSRC, SRC, "-L", LIBDIR, "-lquaru_hip",
                    f"-Wl,-rpath,{LIBDIR}", "-o", out], check=True, capture_output=True,
                   timeout=600)
    lib = ctypes.CDLL(out)
    lib.emu_run2.restype = ctypes.c_int
    lib.emu_run2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int] + \
        [ctypes.c_void_p] * 6 + [ctypes.c_int] + [ctypes.c_void_p] * 2 + \
        [ctypes.c_char_p, ctypes.c_int]
    n, nb, layers = 13, 3, 3
    B = 1 << nb
    rng = np.random.default_rng(13)
    mats, tg, ks, bms, per_state = [], [], [], [], [[] for _ in range(B)]
    for _ in range(layers):
        for axis in ("ry", "rz"):
            for q in range(n):
                th = rng.uniform(0, 2 * np.pi, B)
                ms = [getattr(ops, axis)(float(t), q).matrix() for t in th]
                mats.append(np.concatenate([m.reshape(-1) for m in ms]))
                tg.append([q])
                ks.append(1)
                bms.append(B - 1)
                for b in range(B):
                    per_state[b].append((ms[b], [q]))
        for q in range(n):
            g = ops.cz(q, (q + 1) % n)
            mats.append(g.matrix().reshape(-1))
            tg.append(g.targets())
            ks.append(2)
            bms.append(0)
            for b in range(B):
                per_state[b].append((g.matrix(), g.targets()))
    psi = np.zeros(1 << (n + nb), dtype=np.complex128)
    for b in range(B):
        psi[b << n] = 1.0
    m_all = np.ascontiguousarray(np.concatenate(mats), dtype=np.complex128)
    t_all = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.uint64) for t in tg]))
    ks_a = np.ascontiguousarray(ks, dtype=np.int32)
    bm_a = np.ascontiguousarray(bms, dtype=np.uint32)
    for dtype in ("c128", "c64"):
        st = psi.copy()
        stats = np.zeros(4, dtype=np.uint64)
        err = ctypes.create_string_buffer(512)
        rc = lib.emu_run2(n + nb, 1 if dtype == "c64" else 0, st.ctypes.data, len(ks),
                          m_all.ctypes.data, t_all.ctypes.data, ks_a.ctypes.data, None, None,
                          bm_a.ctypes.data, n, stats.ctypes.data, None, err, 512)
        assert rc == 0, err.value.decode()
        assert stats[0] >= 1 and stats[1] == 0, stats  # sweeps only
        # every RZ(q) folds into the RY(q) before it (state by state)
        assert stats[3] == layers * n, stats
        for b in range(B):
            ref = np.zeros(1 << n, dtype=np.complex128)
            ref[0] = 1.0
            for m, t in per_state[b]:
                ref = oracle.apply_fast(ref, m, t)
            assert np.max(np.abs(st[b << n:(b + 1) << n] - ref)) < TOL[dtype], (dtype, b)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_merge_across_a_changed_pattern_qubit(emu, dtype):
    """U(3); CX(4->3); RZ(4); H(4); CX(4->3) at 14 qubits: RZ and H merge into
    one plain op on qubit 4 that is no longer diagonal, so the second CX must
    stay after it.  (Folding it back into the first CX once gave an error of
    2.7e-2.)  Then random chains of that shape."""
    n = 14
    rng = np.random.default_rng(2718)
    opl = [ops.Operation(random_unitary(rng, 2), [3]), ops.cnot(4, 3), ops.rz(0.8, 4),
           ops.hadamard(4), ops.cnot(4, 3)]
    psi = random_state(rng, n)
    got, st = emu(n, dtype, psi, opl)
    assert st["sweeps"] == 1 and st["merged"] >= 2, st
    assert np.max(np.abs(got - oracle_run(psi.copy(), opl))) < TOL[dtype]
    for n in (11, 14, 16):
        opl = merge_chain_ops(rng, n, 300)
        psi = random_state(rng, n)
        got, st = emu(n, dtype, psi, opl)
        assert st["merged"] > 20
        assert np.max(np.abs(got - oracle_run(psi.copy(), opl))) < TOL[dtype]


# Dispatch codes of the lane-target ops, as qs_sweep_kernel.h numbers them:
# per one-qubit form 5 register slots then 6 lane bits; the one-register-
# pattern multiplexed form at CODE_MUXRL + slot * 6 + lane bit.
_LANE_CODES = 5 + 6
CODE_U1 = 1
CODE_C1 = CODE_U1 + _LANE_CODES
CODE_X1 = CODE_C1 + _LANE_CODES
CODE_MUXL = CODE_X1 + _LANE_CODES + 2 + 25
CODE_MUXRL = CODE_MUXL + _LANE_CODES + 5


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_gpu_workloads_reach_every_lane_bit(emu, dtype):
    """The sweep kernel compiles each lane-target op once per lane bit (DPP
    moves for bits 0-3, permlane swaps for 4 and 5).  The merge-stress
    workloads of the GPU parity test (tests/test_gpu_register.py) must reach
    every lane bit of every one-qubit form, and lane bits 4 and 5 of the
    one-register-pattern form, so the GPU tests run each of those paths."""
    hist = np.zeros(128, dtype=np.uint64)
    for n in (11, 14):
        rng = np.random.default_rng(77 + n)
        _, st = emu(n, dtype, None, merge_stress_ops(rng, n, 400))
        hist += st["hist"]
    for name, base in (("U1", CODE_U1), ("C1", CODE_C1), ("X1", CODE_X1), ("MUXL", CODE_MUXL)):
        missing = [k for k in range(6) if hist[base + 5 + k] == 0]
        assert not missing, f"{name}: lane bits {missing} never planned"
    for k in (4, 5):
        assert sum(int(hist[CODE_MUXRL + j * 6 + k]) for j in range(5)) > 0, k


@pytest.mark.parametrize("n,layers,triples", [(12, 4, 0), (14, 3, 4), (16, 2, 8), (17, 2, 8)])
def test_random_circuit_plans(emu, n, layers, triples):
    w = circuits.random_circuit(n, layers, seed=n * 31 + layers, triples=triples, measure=False)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    got, st = emu(n, "c128", psi, w.ops)
    ref = oracle_run(psi.copy(), w.ops)
    assert st["sweeps"] > 0
    assert np.max(np.abs(got - ref)) < TOL["c128"]


@pytest.mark.parametrize("n", [10, 13])
def test_gate_zoo_plans(emu, n):
    rng = np.random.default_rng(300 + n)
    psi = random_state(rng, n)
    zoo = gate_zoo(rng, n) + gate_zoo(rng, n)
    got, _ = emu(n, "c128", psi, zoo)
    assert np.max(np.abs(got - oracle_run(psi.copy(), zoo))) < TOL["c128"]


def monomial_gate(rng, targets):
    """A permutation with unit factors on `targets` (the form to_quantum_gate
    builds, with phases added)."""
    d = 1 << len(targets)
    m = np.zeros((d, d), dtype=np.complex128)
    m[np.arange(d), rng.permutation(d)] = np.exp(1j * rng.uniform(0, 2 * np.pi, d))
    return ops.Operation(m, list(targets))


def test_permutation_gates_plan(emu):
    """Permutation gates on more than 5 qubits (the gather form) between
    sweeps: classified tables and ordering match the oracle."""
    rng = np.random.default_rng(5)
    n = 11
    opl = []
    for k in (6, 7, 8):
        t = [int(x) for x in rng.permutation(n)[:k]]
        opl.append(monomial_gate(rng, t))
        opl.append(ops.Operation(random_unitary(rng, 2), [t[0]]))
        opl.append(ops.to_controlled(monomial_gate(rng, t[1:]), t[0]))
    opl.append(ops.to_quantum_gate(lambda x: (x * 7) % 64 if x < 64 else x, list(range(1, 8))))
    psi = random_state(rng, n)
    got, st = emu(n, "c128", psi, opl)
    assert st["singles"] >= 7
    assert np.max(np.abs(got - oracle_run(psi.copy(), opl))) < TOL["c128"]


def test_plan_sizes_of_the_benchmark_workloads(emu):
    """Plan-only runs of the 30-qubit random circuit and the 32-qubit Grover
    workload (no state): the sweep counts the planner reaches, as guards
    against planning regressions (holding gates back in front of a
    register-wide phase once took Grover from 100 sweeps to 268).  Registers
    of 12 (complex128) or 13 (complex64) or more qubits use workgroup tiles
    (one wave bit open to dense gates)."""
    w = circuits.random_circuit(30, 20, seed=2025, triples=8, measure=False)
    _, st = emu(30, "c128", None, w.ops)
    assert st["sweeps"] <= 74 and st["merged"] >= 550, st
    _, st = emu(30, "c64", None, w.ops)
    assert st["sweeps"] <= 82, st
    _, opl = circuits.grover_operations(32, 12, seed=4, fused=False)
    _, st = emu(32, "c128", None, opl)
    assert st["sweeps"] <= 80, st
