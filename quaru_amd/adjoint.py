"""Adjoint-method gradients of <psi(theta)|H|psi(theta)> on the GPU.

Thin wrapper over the ``qs_pcircuit_*`` / ``qs_adjoint_*`` C ABI:

* :class:`ParamCircuit` -- fixed gates (k <= 2) and rotations
  exp(-i theta_p sigma/2) about X, Y or Z, with layer marks.  Rotation
  conventions follow the OpenQASM U gate of the reference
  (src/openqasm.rs:145-163): RY(t) = u(t, 0, 0), RZ(t) = u(0, 0, t).
* :class:`Expectation` -- ``forward(thetas)`` returns the batch of energies
  and keeps the final states on the device; ``backward()`` runs one reverse
  sweep and returns dE/dtheta for the whole batch (the Function-style
  forward/backward pair).
* :class:`ExpectationFunction` -- a ``torch.autograd.Function`` over the
  same pair, so the engine plugs into PyTorch-ROCm autograd.

H is a sum of Pauli-Z strings, given as (masks, coeffs): term t contributes
coeffs[t] * (-1)^popcount(x & masks[t]) on basis state x.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .register import _check, _dtype_code, _sizes, _vp

AXIS = {"x": 0, "y": 1, "z": 2}


class ParamCircuit:
    def __init__(self, n: int):
        h = ctypes.c_void_p()
        _check(_lib.load().qs_pcircuit_new(int(n), ctypes.byref(h)))
        self._h = h
        self.n = int(n)

    def close(self):
        if self._h is not None and self._h.value:
            _lib.load().qs_pcircuit_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gate(self, matrix, targets: Sequence[int]) -> "ParamCircuit":
        m = np.ascontiguousarray(matrix, dtype=np.complex128)
        t = _sizes(targets)
        _check(_lib.load().qs_pcircuit_gate(self._h, _vp(m), _vp(t), t.size))
        return self

    def rotation(self, axis: str, target: int, param: int) -> "ParamCircuit":
        _check(_lib.load().qs_pcircuit_rotation(self._h, AXIS[axis.lower()], int(target),
                                                int(param)))
        return self

    def layer(self) -> "ParamCircuit":
        _check(_lib.load().qs_pcircuit_layer(self._h))
        return self

    @property
    def nparams(self) -> int:
        return int(_lib.load().qs_pcircuit_nparams(self._h))

    @staticmethod
    def from_ansatz(ansatz) -> "ParamCircuit":
        """Build from :class:`quaru_amd.circuits.Ansatz`."""
        from .operation import cz
        pc = ParamCircuit(ansatz.n)
        ends = set(ansatz.layer_ends)
        for i, op in enumerate(ansatz.ops):
            if op[0] in ("rx", "ry", "rz"):
                pc.rotation(op[0][1], op[1], op[2])
            elif op[0] == "cz":
                g = cz(op[1], op[2])
                pc.gate(g.matrix(), g.targets())
            else:
                raise ValueError(f"unknown ansatz op {op[0]}")
            if i + 1 in ends:
                pc.layer()
        return pc


class AdjointRun:
    """The device state one forward pass leaves for its backward: the final
    states (and checkpoints) of one batch.  ``backward()`` runs the reverse
    sweep once and frees them; every forward gets its own run, so several
    forwards may be pending at once."""

    def __init__(self, ctx: ctypes.c_void_p, batch: int, nparams: int):
        self._ctx = ctx
        self.batch = batch
        self.nparams = nparams
        self.reverse_stats = None  # set by backward()

    def close(self):
        if self._ctx is not None and self._ctx.value:
            _lib.load().qs_adjoint_free(self._ctx)
        self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        """Sweeps and per-gate kernels of this run's fused forward pass."""
        if self._ctx is None:
            raise RuntimeError("the run is closed")
        sw, kk = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(_lib.load().qs_adjoint_stats(self._ctx, ctypes.byref(sw), ctypes.byref(kk)))
        return {"sweeps": sw.value, "kernels": kk.value}

    def backward(self) -> np.ndarray:
        """dE_b / dtheta_{b,p} as a (batch, nparams) array."""
        if self._ctx is None:
            raise RuntimeError("backward() already ran for this forward pass")
        g = np.zeros((self.batch, self.nparams), dtype=np.float64)
        try:
            lib = _lib.load()
            _check(lib.qs_adjoint_backward(self._ctx, _vp(g)))
            bl, rs = ctypes.c_uint64(0), ctypes.c_uint64(0)
            _check(lib.qs_adjoint_reverse_stats(self._ctx, ctypes.byref(bl), ctypes.byref(rs)))
            # tiled passes over psi and lambda, checkpoint restores
            self.reverse_stats = {"blocks": bl.value, "restores": rs.value}
        finally:
            self.close()
        return g


class Expectation:
    """Energies and adjoint gradients of a batch of parameter sets."""

    def __init__(self, circuit: ParamCircuit, masks, coeffs, dtype="c128",
                 checkpoint_every: int = 0):
        self.circuit = circuit
        self.masks = np.ascontiguousarray(np.asarray(masks, dtype=np.uint64).reshape(-1))
        self.coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64).reshape(-1))
        if self.masks.size != self.coeffs.size:
            raise ValueError("masks and coeffs differ in length")
        self.dtype = _dtype_code(dtype)
        self.checkpoint_every = int(checkpoint_every)
        self._run: Optional[AdjointRun] = None

    def _thetas(self, thetas) -> np.ndarray:
        th = np.ascontiguousarray(np.asarray(thetas, dtype=np.float64))
        if th.ndim == 1:
            th = th.reshape(1, -1)
        if th.ndim != 2 or th.shape[1] != self.circuit.nparams:
            raise ValueError(f"expected (batch, {self.circuit.nparams}) parameters, "
                             f"got shape {np.shape(thetas)}")
        return np.ascontiguousarray(th)

    def energies(self, thetas) -> np.ndarray:
        """Forward only; nothing is kept on the device."""
        th = self._thetas(thetas)
        out = np.zeros(th.shape[0], dtype=np.float64)
        _check(_lib.load().qs_adjoint_forward(self.circuit._h, _vp(th), th.shape[0],
                                              _vp(self.masks), _vp(self.coeffs), self.masks.size,
                                              self.dtype, 0, None, _vp(out)))
        return out

    def record(self, thetas):
        """Forward pass that returns (energies, AdjointRun): the run holds
        this pass's device state until its own backward()."""
        th = self._thetas(thetas)
        out = np.zeros(th.shape[0], dtype=np.float64)
        ctx = ctypes.c_void_p()
        _check(_lib.load().qs_adjoint_forward(self.circuit._h, _vp(th), th.shape[0],
                                              _vp(self.masks), _vp(self.coeffs), self.masks.size,
                                              self.dtype, self.checkpoint_every,
                                              ctypes.byref(ctx), _vp(out)))
        return out, AdjointRun(ctx, th.shape[0], self.circuit.nparams)

    def forward(self, thetas) -> np.ndarray:
        """Function-style forward: energies of the batch; the state for the
        next backward() is kept (a new forward replaces it)."""
        if self._run is not None:
            self._run.close()
        out, self._run = self.record(thetas)
        return out

    def backward(self) -> np.ndarray:
        """Gradients of the last forward() as a (batch, nparams) array."""
        if self._run is None:
            raise RuntimeError("backward() needs a forward() first")
        run, self._run = self._run, None
        return run.backward()


def _torch_function():
    import torch

    class ExpectationFunction(torch.autograd.Function):
        """``E = ExpectationFunction.apply(thetas, expectation)``: thetas is a
        (batch, nparams) or (nparams,) float64 tensor; E has shape (batch,)
        (one entry for a 1-D thetas).  Each application keeps its own device
        state, so several applications may share one Expectation before a
        single backward."""

        @staticmethod
        def forward(ctx, thetas, expectation: Expectation):
            th = thetas.detach().to("cpu", torch.float64).numpy()
            e, run = expectation.record(th)
            ctx.run = run
            ctx.theta_shape = thetas.shape
            ctx.src_device = thetas.device
            return torch.from_numpy(e).to(thetas.device)

        @staticmethod
        def backward(ctx, grad_out):
            g = torch.from_numpy(ctx.run.backward()).to(ctx.src_device)
            ctx.run = None
            grad = grad_out.to(g.dtype).unsqueeze(1) * g
            return grad.reshape(ctx.theta_shape), None

    return ExpectationFunction


_EF = None


def expectation_function():
    """The torch.autograd.Function class (imports torch on first use)."""
    global _EF
    if _EF is None:
        _EF = _torch_function()
    return _EF
