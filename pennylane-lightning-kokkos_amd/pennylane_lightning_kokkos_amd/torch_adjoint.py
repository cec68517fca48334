"""Differentiable expectation values on the gfx950 engine.

`AdjointExpval` is a torch.autograd.Function: forward prepares |0...0>,
applies the circuit and returns the expectation value of every observable;
backward returns dy^T J from one fused adjoint sweep of the single observable
H = sum_j dy_j O_j, the construction of the reference's `vjp`
(pennylane_lightning_kokkos/lightning_kokkos.py:689-728). torch is only the
autograd and tensor edge: states live in the engine's HBM buffers.

Circuits are lists of operations (name, wires, params, inverse[, matrix]);
`Rot` is expanded into RZ, RY, RZ as the reference serializer does
(`_serialize.py:283-284`, `rotFunctor` GateFunctors.hpp:3148-3163).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

from . import _capi


@dataclass
class Op:
    name: str
    wires: List[int]
    params: List[float] = field(default_factory=list)
    inverse: bool = False
    matrix: object = None


def expand_ops(ops: Sequence[Op]):
    """Serialized op list and, per serialized parameter, the index of the
    circuit parameter it comes from (Rot -> RZ(phi), RY(theta), RZ(omega))."""
    names, params, wires, inverses, mats, origin = [], [], [], [], [], []
    k = 0
    for op in ops:
        if op.name == "Rot":
            phi, theta, omega = op.params
            # Rot = RZ(omega) RY(theta) RZ(phi); its adjoint applies the
            # inverted factors in the opposite order.
            seq = [("RZ", phi), ("RY", theta), ("RZ", omega)]
            if op.inverse:
                seq = seq[::-1]
            for j, (g, a) in enumerate(seq):
                names.append(g)
                params.append([float(a)])
                wires.append(list(op.wires))
                inverses.append(bool(op.inverse))
                mats.append([])
                origin.append(k + (j if not op.inverse else 2 - j))
            k += 3
            continue
        names.append(op.name)
        params.append([float(p) for p in op.params])
        wires.append(list(op.wires))
        inverses.append(bool(op.inverse))
        mats.append([] if op.matrix is None else np.asarray(op.matrix, dtype=np.complex128).reshape(-1))
        for _ in op.params:
            origin.append(k)
            k += 1
    return (names, params, wires, inverses, mats), origin, k


class Circuit:
    """A parametrized circuit with observables, bound to one engine state."""

    def __init__(self, num_qubits: int, ops: Sequence[Op], observables: Sequence[tuple],
                 dtype=np.complex128, device_id: int = 0):
        self.n = int(num_qubits)
        self.ops = list(ops)
        self.obs_desc = list(observables)
        self.dtype = np.dtype(dtype)
        self.device_id = device_id
        self.state = _capi.State(self.n, self.dtype, device_id)
        self.num_params = sum(3 if o.name == "Rot" else len(o.params) for o in self.ops)

    def _with_params(self, values: np.ndarray) -> List[Op]:
        out, k = [], 0
        for op in self.ops:
            m = 3 if op.name == "Rot" else len(op.params)
            out.append(Op(op.name, op.wires, list(values[k:k + m]), op.inverse, op.matrix))
            k += m
        return out

    def _obs(self, desc):
        kind = desc[0]
        if kind == "named":
            return _capi.Obs.named(desc[1], desc[2])
        if kind == "herm":
            return _capi.Obs.hermitian(desc[1], desc[2])
        if kind == "tensor":
            return _capi.Obs.tensor([self._obs(f) for f in desc[1]])
        if kind == "ham":
            return _capi.Obs.hamiltonian(desc[1], [self._obs(t) for t in desc[2]])
        if kind == "sparse":
            return _capi.Obs.sparse(desc[1], desc[2], desc[3], desc[4])
        raise ValueError(f"unknown observable {kind}")

    def expvals(self, values: np.ndarray) -> np.ndarray:
        ser, _, _ = expand_ops(self._with_params(values))
        self.state.reset()
        self.state.apply_ops(_capi.Ops(*ser))
        return np.array([self.state.expval(self._obs(d)) for d in self.obs_desc])

    def jacobian(self, values: np.ndarray, weights: np.ndarray = None) -> np.ndarray:
        """d<O_j>/d values (n_obs x num_params), or weights^T J when weights
        are given (one sweep of sum_j weights_j O_j)."""
        ser, origin, k = expand_ops(self._with_params(values))
        npar = sum(1 for p in ser[1] if p)
        if weights is None:
            obs = [self._obs(d) for d in self.obs_desc]
        else:
            obs = [_capi.Obs.hamiltonian(np.asarray(weights, dtype=np.float64),
                                         [self._obs(d) for d in self.obs_desc])]
        self.state.reset()
        jac = _capi.adjoint_jacobian(self.state, obs, _capi.Ops(*ser), list(range(npar)), True)
        out = np.zeros((jac.shape[0], k))
        for col, src in enumerate(origin):
            out[:, src] += jac[:, col]
        return out


def _torch():
    import torch

    return torch


def make_function():
    torch = _torch()

    class AdjointExpval(torch.autograd.Function):
        """expvals = AdjointExpval.apply(params, circuit)."""

        @staticmethod
        def forward(ctx, params, circuit: Circuit):
            values = params.detach().to("cpu", torch.float64).numpy()
            ctx.circuit = circuit
            ctx.values = values
            ctx.device = params.device
            ctx.dtype = params.dtype
            return torch.as_tensor(circuit.expvals(values), dtype=params.dtype, device=params.device)

        @staticmethod
        def backward(ctx, dy):
            w = dy.detach().to("cpu", torch.float64).numpy()
            if not np.any(w):
                return torch.zeros(len(ctx.values), dtype=ctx.dtype, device=ctx.device), None
            g = ctx.circuit.jacobian(ctx.values, weights=w)[0]
            return torch.as_tensor(g, dtype=ctx.dtype, device=ctx.device), None

    return AdjointExpval
