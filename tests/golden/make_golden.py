"""Write the golden fixtures under tests/golden/ (numpy only).

The reference (Rust) cannot be built or run here, and ships no stored
vectors.  These fixtures are computed with numpy by an algorithm independent
of the C oracle's gather/scatter loop: each gate is a tensor contraction of
the state reshaped to [2]*n (qubit q = axis n-1-q, src/register.rs:41-50)
with the gate reshaped to [2]*2k (matrix-index bit j <-> targets[j],
src/register.rs:165-176).  Measurement follows try_measure
(src/register.rs:297-342): res = x > prob_0, collapse, divide by sqrt.

Fixtures:
  c1_12q.npz    the 12-qubit random circuit: 20 layers of Haar 2x2 on
                every qubit + cnot(q, q+1) ladder, seed 2024; final
                pre-measure amplitudes, the uniforms, and the index sampled
                by measuring qubits 0..11 in order.
  c3_mini.npz   the Toffoli/CCZ layer structure at 10 qubits, 3 layers, 4
                Toffoli/CCZ per layer, seed 7 (Toffoli / CCZ coverage).
  adj_4q.npz    4-qubit, 2-layer RY/RZ + CZ-ring ansatz: energy and exact
                parameter-shift gradients of H = sum Z_i + 0.5 sum Z_i Z_i+1
                for 3 seeded parameter sets.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from quaru_amd import circuits  # noqa: E402
from quaru_amd import operation as ops  # noqa: E402


def apply_np(psi: np.ndarray, m: np.ndarray, targets, n: int) -> np.ndarray:
    k = len(targets)
    t = psi.reshape([2] * n)
    mt = np.asarray(m, dtype=np.complex128).reshape([2] * (2 * k))
    col_axes = [k + (k - 1 - j) for j in range(k)]
    st_axes = [n - 1 - targets[j] for j in range(k)]
    out = np.tensordot(mt, t, axes=(col_axes, st_axes))
    out = np.moveaxis(out, [k - 1 - j for j in range(k)], st_axes)
    return np.ascontiguousarray(out).reshape(-1)


def measure_np(psi: np.ndarray, target: int, x: float):
    idx = np.arange(psi.size)
    one = ((idx >> target) & 1).astype(bool)
    p = (psi.conj() * psi).real
    p0 = float(np.sum(p[~one]))
    p1 = float(np.sum(p[one]))
    res = x > p0
    total = p1 if res else p0
    out = psi.copy()
    out[one != res] = 0.0
    out[one == res] /= math.sqrt(total)
    return res, out


def run_workload(w: circuits.Workload) -> np.ndarray:
    psi = np.zeros(1 << w.n, dtype=np.complex128)
    psi[0] = 1.0
    for op in w.ops:
        psi = apply_np(psi, op.matrix(), op.targets(), w.n)
    return psi


def c1():
    w = circuits.random_circuit(12, 20, seed=2024)
    psi = run_workload(w)
    state = psi.copy()
    idx = 0
    for q in range(w.n):
        res, state = measure_np(state, q, w.uniforms[q])
        idx |= int(res) << q
    np.savez(os.path.join(HERE, "c1_12q.npz"), n=12, layers=20, seed=2024, amplitudes=psi,
             uniforms=np.asarray(w.uniforms), index=idx, collapsed=state)
    print("c1_12q: index", idx, "norm", np.vdot(psi, psi).real)


def c3_mini():
    w = circuits.random_circuit(10, 3, seed=7, triples=4, measure=False)
    psi = run_workload(w)
    np.savez(os.path.join(HERE, "c3_mini.npz"), n=10, layers=3, seed=7, triples=4,
             amplitudes=psi)
    print("c3_mini: norm", np.vdot(psi, psi).real)


def ansatz_energy(n, ans, theta, masks, coeffs):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    for op in ans.ops:
        if op[0] == "ry":
            g = ops.ry(theta[op[2]], op[1])
        elif op[0] == "rz":
            g = ops.rz(theta[op[2]], op[1])
        else:
            g = ops.cz(op[1], op[2])
        psi = apply_np(psi, g.matrix(), g.targets(), n)
    x = np.arange(1 << n, dtype=np.uint64)
    h = np.zeros(1 << n)
    for mk, c in zip(masks, coeffs):
        par = np.zeros(1 << n, dtype=np.uint64)
        v = x & np.uint64(mk)
        while np.any(v):
            par ^= v & np.uint64(1)
            v = v >> np.uint64(1)
        h += np.where(par == 1, -c, c)
    return float(np.sum((psi.conj() * psi).real * h))


def adj_4q():
    n, layers = 4, 2
    ans = circuits.hardware_efficient_ansatz(n, layers)
    masks, coeffs = circuits.zz_chain_observable(n)
    thetas = circuits.seeded_thetas(3, ans.nparams, seed=11)
    energies, grads = [], []
    for th in thetas:
        energies.append(ansatz_energy(n, ans, th, masks, coeffs))
        g = np.zeros(ans.nparams)
        for p in range(ans.nparams):
            tp, tm = th.copy(), th.copy()
            tp[p] += math.pi / 2
            tm[p] -= math.pi / 2
            g[p] = 0.5 * (ansatz_energy(n, ans, tp, masks, coeffs)
                          - ansatz_energy(n, ans, tm, masks, coeffs))
        grads.append(g)
    np.savez(os.path.join(HERE, "adj_4q.npz"), n=n, layers=layers, thetas=thetas,
             energies=np.asarray(energies), grads=np.asarray(grads), masks=masks,
             coeffs=coeffs)
    print("adj_4q: energies", energies)


if __name__ == "__main__":
    c1()
    c3_mini()
    adj_4q()
