"""Batched forward + adjoint of many small circuits in one persistent launch."""
from __future__ import annotations

from typing import Sequence

import numpy as np

from . import _native as N
from .circuit import _dtype_code

PARAMETRIC = (N.GATE_RX, N.GATE_RY, N.GATE_RZ)


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
    count = len(circuits)
    if count == 0:
        return np.zeros(0), []
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
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    ctrls = np.ascontiguousarray(ctrls, dtype=np.int32)
    tgts = np.ascontiguousarray(tgts, dtype=np.int32)
    params = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    coef = np.ascontiguousarray(np.concatenate(coefs))
    energy = np.empty(count)
    grad = np.empty(max(1, p_total))
    N.call("dmv_batch_forward_adjoint", count, _dtype_code(dtype), N.iptr(n), N.iptr(gate_offset), N.iptr(kinds),
           N.iptr(ctrls), N.iptr(tgts), N.dptr(params), N.iptr(coef_offset), N.dptr(coef), N.iptr(param_offset),
           N.dptr(energy), N.dptr(grad))
    grads = [grad[param_offset[i]: param_offset[i] + n_params[i]].copy() for i in range(count)]
    return energy, grads
