"""Batched forward + adjoint of many small circuits in one persistent launch."""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _native as N
from .circuit import _dtype_code

PARAMETRIC = (N.GATE_RX, N.GATE_RY, N.GATE_RZ)


class PackedBatch:
    """Circuits of a batch in the flat arrays of dmv_batch_forward_adjoint
    (include/damavand_gpu.h): per-circuit qubit counts, gate / coefficient /
    parameter offsets, and the gates and coefficients of all circuits."""

    def __init__(self, n, gate_offset, kinds, ctrls, tgts, params, coef_offset, coef, param_offset, n_params):
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.gate_offset = np.ascontiguousarray(gate_offset, dtype=np.int32)
        self.kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        self.ctrls = np.ascontiguousarray(ctrls, dtype=np.int32)
        self.tgts = np.ascontiguousarray(tgts, dtype=np.int32)
        self.params = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.coef_offset = np.ascontiguousarray(coef_offset, dtype=np.int32)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64)
        self.param_offset = np.ascontiguousarray(param_offset, dtype=np.int32)
        self.n_params = np.ascontiguousarray(n_params, dtype=np.int64)

    @property
    def count(self) -> int:
        return int(self.n.size)

    @property
    def total_params(self) -> int:
        return int(self.n_params.sum())

    def parameters(self) -> np.ndarray:
        """The RX/RY/RZ angles of all circuits, circuit after circuit in gate
        order: the order of the gradients."""
        return self.params.reshape(-1, 3)[self._param_rows(), 0].copy()

    def with_parameters(self, theta) -> "PackedBatch":
        """A copy whose RX/RY/RZ angles are ``theta`` (same order as
        ``parameters()``); gates, coefficients and offsets are shared."""
        theta = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        rows = self._param_rows()
        if theta.size != rows.size:
            raise ValueError(f"need {rows.size} angles, got {theta.size}")
        params = self.params.reshape(-1, 3).copy()
        params[rows, 0] = theta
        out = object.__new__(PackedBatch)
        out.__dict__.update(self.__dict__)
        out.params = params.reshape(-1)
        return out

    def _param_rows(self) -> np.ndarray:
        return np.flatnonzero(np.isin(self.kinds, PARAMETRIC))


def pack_circuits(circuits: Sequence) -> PackedBatch:
    """``circuits``: a sequence of ``(n, gates, a, b)`` (see batch_forward_adjoint)."""
    count = len(circuits)
    n = np.zeros(count, dtype=np.int32)
    gate_offset = np.zeros(count + 1, dtype=np.int32)
    coef_offset = np.zeros(count, dtype=np.int32)
    param_offset = np.zeros(count, dtype=np.int32)
    kinds, ctrls, tgts, params, coefs = [], [], [], [], []
    n_params = []
    g_total = c_total = p_total = 0
    for i, (ni, gates, a, b) in enumerate(circuits):
        n[i] = ni
        a = np.asarray(a, dtype=np.float64).reshape(-1)
        b = np.asarray(b, dtype=np.float64).reshape(-1)
        if a.size != ni or b.size != ni - 1:
            raise ValueError(f"circuit {i}: need {ni} Z and {ni - 1} ZZ coefficients")
        gate_offset[i] = g_total
        coef_offset[i] = c_total
        param_offset[i] = p_total
        np_i = 0
        for kind, control, target, prm in gates:
            kinds.append(kind)
            ctrls.append(control)
            tgts.append(target)
            row = [0.0, 0.0, 0.0]
            row[: len(prm)] = list(prm)
            params.append(row)
            if kind in PARAMETRIC:
                np_i += 1
        g_total += len(gates)
        coefs.append(a)
        coefs.append(b)
        c_total += 2 * ni - 1
        p_total += np_i
        n_params.append(np_i)
    gate_offset[count] = g_total
    return PackedBatch(n, gate_offset, kinds, ctrls, tgts, np.asarray(params, dtype=np.float64).reshape(-1),
                       coef_offset, np.concatenate(coefs) if coefs else np.zeros(0), param_offset, n_params)


def batch_forward_adjoint_packed(batch: PackedBatch, dtype="complex128"):
    """One persistent launch over a packed batch: returns (energies, flat
    gradient array in the order of ``batch.param_offset``)."""
    energy = np.empty(batch.count)
    grad = np.empty(max(1, batch.total_params))
    N.call("dmv_batch_forward_adjoint", batch.count, _dtype_code(dtype), N.iptr(batch.n), N.iptr(batch.gate_offset),
           N.iptr(batch.kinds), N.iptr(batch.ctrls), N.iptr(batch.tgts), N.dptr(batch.params),
           N.iptr(batch.coef_offset), N.dptr(batch.coef), N.iptr(batch.param_offset), N.dptr(energy), N.dptr(grad))
    return energy, grad[: batch.total_params]


def batch_autotune(batch: PackedBatch, dtype="complex128") -> int:
    """Times the batched kernel's launch shapes on this packed batch and keeps
    the fastest for batches whose largest circuit has as many qubits (in this
    process and in the file named by DMV_TUNE_CACHE).  Returns the chosen tile
    qubits."""
    k = ctypes.c_int(0)
    N.call("dmv_batch_autotune", batch.count, _dtype_code(dtype), N.iptr(batch.n), N.iptr(batch.gate_offset),
           N.iptr(batch.kinds), N.iptr(batch.ctrls), N.iptr(batch.tgts), N.dptr(batch.params),
           N.iptr(batch.coef_offset), N.dptr(batch.coef), N.iptr(batch.param_offset), ctypes.byref(k))
    return k.value


def batch_forward_adjoint(circuits: Sequence, dtype="complex128"):
    """Runs every circuit from |0>, evaluates its chain observable and its
    adjoint gradient, all in one persistent kernel launch.

    ``circuits`` is a sequence of ``(n, gates, a, b)``: ``gates`` a list of
    ``(kind, control, target, params)``, ``a`` the n Z coefficients and ``b``
    the n-1 nearest-neighbour ZZ coefficients of
    H = sum_q a_q Z_q + sum_{q<n-1} b_q Z_q Z_{q+1}.

    Returns ``(energies, grads)``: an array of <H> per circuit and a list of
    gradient arrays (one entry per RX/RY/RZ gate, in insertion order).
    """
    if len(circuits) == 0:
        return np.zeros(0), []
    batch = pack_circuits(circuits)
    energy, grad = batch_forward_adjoint_packed(batch, dtype)
    grads = [grad[batch.param_offset[i]: batch.param_offset[i] + batch.n_params[i]].copy()
             for i in range(batch.count)]
    return energy, grads
