"""Dense np.kron simulator used to pin the C oracle -- TEST INFRASTRUCTURE ONLY.

Restates the reference's ``brute_force`` method: the full 2^n x 2^n operator of
a gate is the Kronecker product of per-qubit 2x2 blocks, listed from the
highest qubit down to qubit 0 (src/utils.rs:187-207, so qubit q is bit q of the
amplitude index), and a CNOT is P0(c) (x) I + P1(c) (x) X(t)
(src/utils.rs:208-245, src/qubit_backend/circuit_brute_force.rs:9-12).
Only for n <= 10.
"""
from __future__ import annotations

import numpy as np

_I = np.eye(2, dtype=np.complex128)
_P0 = np.array([[1, 0], [0, 0]], dtype=np.complex128)
_P1 = np.array([[0, 0], [0, 1]], dtype=np.complex128)


def _kron_list(mats):
    out = np.array([[1.0 + 0j]])
    for m in mats:
        out = np.kron(out, m)
    return out


def full_operator(n: int, m: np.ndarray, target: int, controls=()) -> np.ndarray:
    """2^n x 2^n operator of a 2x2 matrix on `target` controlled on every qubit in `controls`."""
    controls = tuple(controls)
    if not controls:
        return _kron_list([m if q == target else _I for q in range(n - 1, -1, -1)])
    # sum over control patterns: identity unless every control is 1
    active = _kron_list([m if q == target else (_P1 if q in controls else _I)
                         for q in range(n - 1, -1, -1)])
    proj_all = _kron_list([_P1 if q in controls else _I for q in range(n - 1, -1, -1)])
    return np.eye(1 << n, dtype=np.complex128) - proj_all + active


def run(n: int, ops) -> np.ndarray:
    """ops: [(matrix2x2, target, controls_tuple), ...] applied in order to |0>."""
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    for m, t, cs in ops:
        psi = full_operator(n, np.asarray(m, dtype=np.complex128), t, cs) @ psi
    return psi
