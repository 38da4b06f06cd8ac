"""GPU adjoint gradients against parameter shift (CPU oracle / numpy
goldens), central finite differences on the GPU, and closed forms."""
import math

import numpy as np
import pytest

import oracle
from quaru_amd import circuits
from quaru_amd import operation as ops
from quaru_amd.adjoint import Expectation, ParamCircuit, expectation_function

pytestmark = pytest.mark.gpu


def oracle_energy(ans, theta, masks, coeffs):
    """Forward pass of the ansatz with the CPU oracle (FP64)."""
    s = oracle.new([False] * ans.n)
    for op in ans.ops:
        if op[0] == "ry":
            g = ops.ry(theta[op[2]], op[1])
        elif op[0] == "rz":
            g = ops.rz(theta[op[2]], op[1])
        else:
            g = ops.cz(op[1], op[2])
        s = oracle.apply_fast(s, g.matrix(), g.targets())
    return oracle.expect_zstrings(s, masks, coeffs)


def test_one_qubit_closed_form():
    pc = ParamCircuit(1).rotation("y", 0, 0).layer()
    ex = Expectation(pc, [1], [1.0])
    th = np.array([[0.3], [1.2], [2.9]])
    e = ex.forward(th)
    g = ex.backward()
    assert np.max(np.abs(e - np.cos(th[:, 0]))) < 1e-14
    assert np.max(np.abs(g[:, 0] + np.sin(th[:, 0]))) < 1e-14


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_single_rotation_axes(axis):
    # <Z> after H then R_axis(t): X -> cos t on |0>, Y, Z via an H first
    pc = ParamCircuit(1)
    h = ops.hadamard(0)
    pc.gate(h.matrix(), [0]).rotation(axis, 0, 0).gate(h.matrix(), [0]).layer()
    ex = Expectation(pc, [1], [1.0])
    t = np.array([[0.77]])
    e = ex.forward(t)
    g = ex.backward()
    # finite difference on the same engine
    d = 1e-6
    fd = (ex.energies(t + d) - ex.energies(t - d)) / (2 * d)
    assert abs(g[0, 0] - fd[0]) < 1e-8
    assert np.isfinite(e).all()


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 1e-5)])
def test_golden_4q_parameter_shift(golden, dtype, tol):
    g = golden("adj_4q.npz")
    ans = circuits.hardware_efficient_ansatz(int(g["n"]), int(g["layers"]))
    pc = ParamCircuit.from_ansatz(ans)
    ex = Expectation(pc, g["masks"], g["coeffs"], dtype=dtype)
    e = ex.forward(g["thetas"])
    gr = ex.backward()
    assert np.max(np.abs(e - g["energies"])) < tol
    assert np.max(np.abs(gr - g["grads"])) < tol


@pytest.mark.parametrize("checkpoint", [0, 2])
def test_adjoint_vs_fd_and_parameter_shift(checkpoint):
    n, layers, batch = 8, 3, 5
    ans = circuits.hardware_efficient_ansatz(n, layers)
    masks, coeffs = circuits.zz_chain_observable(n)
    th = circuits.seeded_thetas(batch, ans.nparams, seed=3)
    pc = ParamCircuit.from_ansatz(ans)
    ex = Expectation(pc, masks, coeffs, checkpoint_every=checkpoint)
    e = ex.forward(th)
    g = ex.backward()
    # energies vs the oracle
    for b in range(batch):
        assert abs(e[b] - oracle_energy(ans, th[b], masks, coeffs)) < 1e-12
    # central FD (h = 1e-4, FP64) on the GPU for every parameter of set 0,2
    h = 1e-4
    for b in (0, 2):
        plus = np.repeat(th[b:b + 1], ans.nparams, axis=0)
        minus = plus.copy()
        idx = np.arange(ans.nparams)
        plus[idx, idx] += h
        minus[idx, idx] -= h
        fd = (ex.energies(plus) - ex.energies(minus)) / (2 * h)
        assert np.max(np.abs(fd - g[b])) < 1e-6
    # parameter shift with the CPU oracle on a sample of parameters
    for p in (0, 7, ans.nparams - 1):
        tp, tm = th[1].copy(), th[1].copy()
        tp[p] += math.pi / 2
        tm[p] -= math.pi / 2
        ps = 0.5 * (oracle_energy(ans, tp, masks, coeffs) - oracle_energy(ans, tm, masks, coeffs))
        assert abs(ps - g[1, p]) < 1e-11


def test_c64_adjoint_agrees_with_c128():
    n, layers, batch = 10, 4, 4
    ans = circuits.hardware_efficient_ansatz(n, layers)
    masks, coeffs = circuits.zz_chain_observable(n)
    th = circuits.seeded_thetas(batch, ans.nparams, seed=9)
    pc = ParamCircuit.from_ansatz(ans)
    g = {}
    for dt in ("c128", "c64"):
        ex = Expectation(pc, masks, coeffs, dtype=dt, checkpoint_every=2)
        ex.forward(th)
        g[dt] = ex.backward()
    assert np.max(np.abs(g["c64"] - g["c128"])) < 1e-5


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    n, layers = 6, 2
    ans = circuits.hardware_efficient_ansatz(n, layers)
    masks, coeffs = circuits.zz_chain_observable(n)
    ex = Expectation(ParamCircuit.from_ansatz(ans), masks, coeffs)
    th = torch.tensor(circuits.seeded_thetas(3, ans.nparams, seed=1), requires_grad=True)
    E = expectation_function().apply(th, ex)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    (E * w).sum().backward()
    ex2 = Expectation(ParamCircuit.from_ansatz(ans), masks, coeffs)
    ex2.forward(th.detach().numpy())
    ref = ex2.backward() * w.numpy()[:, None]
    assert np.max(np.abs(th.grad.numpy() - ref)) < 1e-12
