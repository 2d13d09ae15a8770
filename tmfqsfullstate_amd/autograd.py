"""torch.autograd.Function over the device adjoint gradient.

``CircuitExpectation.apply(thetas, circuit)`` returns E(thetas) = <psi|O|psi>
for psi = U(thetas)|init> as a 0-d tensor and, on backward, the adjoint
gradient computed on the GPU by ``tmfqs_hip_adjoint`` (one forward and one
backward sweep, whatever the number of angles).  ``thetas`` replaces the
angles of the circuit's CPHASE/PHASE records in order.  torch only carries the
parameters and the gradient on the host; the state lives in the engine's own
device buffers.
"""

import ctypes

import numpy as np
import torch

from . import C128, OP_CPHASE, OP_PHASE, HipState, Op


class AdjointCircuit:
    """An op-record circuit with a diagonal observable, evaluated on the GPU."""

    def __init__(self, n, ops, count, terms, precision=C128, init_state=0, checkpoint_gates=0, mat_pool=None,
                 qubit_pool=None):
        self.n = n
        self.ops = (Op * max(1, count))()
        ctypes.memmove(self.ops, ops, ctypes.sizeof(Op) * count)
        self.count = count
        self.terms = list(terms)
        self.precision = precision
        self.init_state = init_state
        self.checkpoint_gates = checkpoint_gates
        self.mat_pool = mat_pool
        self.qubit_pool = qubit_pool
        self.params = [i for i in range(count) if ops[i].kind in (OP_CPHASE, OP_PHASE)]

    def evaluate(self, thetas):
        """(E, dE/dthetas) as numpy values."""
        th = np.asarray(thetas, dtype=np.float64).reshape(-1)
        if th.size != len(self.params):
            raise ValueError(f"expected {len(self.params)} angles, got {th.size}")
        for v, i in zip(th, self.params):
            self.ops[i].theta = float(v)
        with HipState(precision=self.precision) as s:
            s.init_basis(self.n, self.init_state)
            return s.adjoint(self.ops, self.count, self.terms, self.checkpoint_gates, self.mat_pool, self.qubit_pool)


class CircuitExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, thetas, circuit):
        e, g = circuit.evaluate(thetas.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(g, dtype=thetas.dtype).reshape(thetas.shape))
        return thetas.new_tensor(e)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None
