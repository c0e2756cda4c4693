"""Writes tests/golden/golden.npz: small fixed vectors that pin the oracle, the
workload generator and (on the GPU) the HIP path.

    python tests/golden/make_golden.py

Contents (all data, no code):

* kat_2q: the reference's only numeric known answer, H(0), H(1), CNOT(0,1) on
  2 qubits -> every amplitude 1/2 (src/qubit_backend/circuit.rs:802-837).
* circ_a_* / circ_b_*: two gate lists with amplitudes computed by the dense
  np.kron simulator (oracle/brute.py, the reference's brute_force method),
  independent of oracle.c.
* config1_angles, config3_angles, config3_a, config3_b: the seeded angles and
  observable coefficients of the 20-qubit CPU workload and of the 24-qubit
  adjoint workload (damavand_amd/workloads.py), so a change of generator shows.
* sampler_p / sampler_u / sampler_idx: inverse-CDF known answers worked out by
  hand for src/utils.rs:258-277 (cum = [0, .125, .375, .5, 1]).
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from damavand_amd import workloads as wl  # noqa: E402
from oracle import brute  # noqa: E402

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.diag([1.0 + 0j, -1.0])
Hm = np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2)


def textbook(kind, p):
    """Gate matrices from their definitions (not from the oracle)."""
    c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
    return {
        wl.GATE_H: Hm, wl.GATE_X: X, wl.GATE_Y: Y, wl.GATE_Z: Z, wl.GATE_CNOT: X, wl.GATE_CZ: Z,
        wl.GATE_RX: np.array([[c, -1j * s], [-1j * s, c]]),
        wl.GATE_RY: np.array([[c, -s], [s, c]]),
        wl.GATE_RZ: np.diag([np.exp(-0.5j * p[0]), np.exp(0.5j * p[0])]),
        wl.GATE_S: np.diag([1, 1j]), wl.GATE_T: np.diag([1, np.exp(0.25j * math.pi)]),
        wl.GATE_U3: np.array([[c, -np.exp(1j * p[2]) * s], [np.exp(1j * p[1]) * s, np.exp(1j * (p[1] + p[2])) * c]]),
    }[kind]


def random_gates(seed, n, count):
    rng = np.random.default_rng(seed)
    kinds = [wl.GATE_H, wl.GATE_X, wl.GATE_Y, wl.GATE_Z, wl.GATE_RX, wl.GATE_RY, wl.GATE_RZ, wl.GATE_S, wl.GATE_T,
             wl.GATE_CNOT, wl.GATE_U3, wl.GATE_CZ]
    out = []
    for _ in range(count):
        k = kinds[rng.integers(len(kinds))]
        t = int(rng.integers(n))
        c = -1
        if k in (wl.GATE_CNOT, wl.GATE_CZ):
            c = int(rng.integers(n - 1))
            c += c >= t
        out.append((k, c, t, tuple(rng.uniform(0, 2 * math.pi, 3))))
    return out


def pack(gates):
    kinds = np.array([g[0] for g in gates], dtype=np.int32)
    ctrl = np.array([g[1] for g in gates], dtype=np.int32)
    tgt = np.array([g[2] for g in gates], dtype=np.int32)
    prm = np.array([list(g[3]) + [0.0] * (3 - len(g[3])) for g in gates], dtype=np.float64)
    return kinds, ctrl, tgt, prm


def amplitudes(n, gates):
    return brute.run(n, [(textbook(k, list(p) + [0, 0, 0]), t, () if c < 0 else (c,)) for k, c, t, p in gates])


def main():
    out = {}
    out["kat_2q"] = np.full(4, 0.5 + 0j)
    for name, n, gates in (("circ_a", 6, random_gates(7, 6, 60)),
                           ("circ_b", 9, wl.layered_circuit(9, 2, 11))):
        k, c, t, p = pack(gates)
        out[name + "_n"] = np.array(n)
        out[name + "_kind"], out[name + "_ctrl"], out[name + "_tgt"], out[name + "_prm"] = k, c, t, p
        out[name + "_amp"] = amplitudes(n, gates)
    g1 = wl.layered_circuit(20, 10, 1001)
    out["config1_angles"] = np.array([g[3][0] for g in g1 if g[0] in wl.PARAMETRIC])
    g3, a, b = wl.ansatz_circuit(24, 20, 1003)
    out["config3_angles"] = np.array([g[3][0] for g in g3 if g[0] in wl.PARAMETRIC])
    out["config3_a"], out["config3_b"] = a, b
    out["sampler_p"] = np.array([0.125, 0.25, 0.125, 0.5])
    out["sampler_u"] = np.array([1e-300, 0.125, 0.1250001, 0.375, 0.4, 0.5, 0.5000001, 1.0])
    out["sampler_idx"] = np.array([0, 0, 1, 1, 2, 2, 3, 3])
    # a 3-qubit state whose amplitudes are exactly 0 or 1/2, so every cumulative
    # sum is exact: cum = [0, .25, .25, .5, .5, .75, .75, .75, 1]
    out["sampler2_amp"] = np.array([0.5, 0, 0.5, 0, 0.5, 0, 0, 0.5], dtype=complex)
    out["sampler2_u"] = np.array([1e-300, 0.25, 0.26, 0.5, 0.5000001, 0.75, 0.76, 1.0])
    out["sampler2_idx"] = np.array([0, 0, 2, 2, 4, 4, 7, 7])
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
