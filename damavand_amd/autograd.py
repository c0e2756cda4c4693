"""torch.autograd bridge: E(theta) = <0|U(theta)^dag H U(theta)|0> with its
gradient from the adjoint kernels of libdamavand-gpu.

    energy = expectation(circuit, theta, a, pairs, b)   # theta: torch tensor
    energy.backward()                                   # theta.grad = dE/dtheta

    energies = batch_expectation(packed, theta)         # many circuits, one launch
    energies.sum().backward()

The forward pass computes the energy and the full gradient in one adjoint
sweep (forward, lambda = H psi, reverse sweep); backward only scales the
stored gradient by the incoming grad_output.
"""
from __future__ import annotations

import numpy as np
import torch

from .batch import PackedBatch, batch_forward_adjoint_packed
from .circuit import Circuit


class AdjointExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, circuit: Circuit, a, pairs, b):
        p = theta.detach().to("cpu", torch.float64).numpy()
        circuit.set_parameters(p)
        e, g = circuit.gradient(a, pairs, b)
        ctx.save_for_backward(torch.as_tensor(np.asarray(g), dtype=theta.dtype, device=theta.device))
        return torch.tensor(e, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        (g,) = ctx.saved_tensors
        return grad_output * g, None, None, None, None


def expectation(circuit: Circuit, theta: torch.Tensor, a=None, pairs=(), b=None) -> torch.Tensor:
    if a is None:
        a = np.zeros(circuit.num_qubits)
    return AdjointExpectation.apply(theta, circuit, a, list(pairs), b)


class BatchedAdjointExpectation(torch.autograd.Function):
    """Energies of every circuit of a packed batch at angles ``theta`` (all
    circuits' RX/RY/RZ angles, in PackedBatch.parameters() order), with the
    gradients from the same persistent launch."""

    @staticmethod
    def forward(ctx, theta, batch: PackedBatch, dtype):
        p = theta.detach().to("cpu", torch.float64).numpy()
        e, g = batch_forward_adjoint_packed(batch.with_parameters(p), dtype)
        owner = np.repeat(np.arange(batch.count), batch.n_params)  # circuit of each angle
        ctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype, device=theta.device),
                              torch.as_tensor(owner, device=theta.device))
        return torch.as_tensor(e, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_output):
        g, owner = ctx.saved_tensors
        return g * grad_output[owner], None, None


def batch_expectation(batch: PackedBatch, theta: torch.Tensor, dtype="complex128") -> torch.Tensor:
    """Per-circuit energies of a packed batch (see batch.pack_circuits) as a
    differentiable function of all their angles."""
    return BatchedAdjointExpectation.apply(theta, batch, dtype)
