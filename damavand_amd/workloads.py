"""Seeded synthetic circuits for the statevector path.

A gate is a tuple ``(kind, control, target, params)`` with ``control = -1`` for
none and ``params`` a tuple of floats (empty for fixed gates).  Kinds are the
``DMV_GATE_*`` values of ``include/damavand_gpu.h``.

Random numbers come from splitmix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014; constants as published by
Vigna), written out here and in ``damavand_amd/csrc/rng.hpp`` so that host and
device draw identical streams.  The k-th output of a stream with seed s is
``mix(s + (k + 1) * 0x9E3779B97F4A7C15)``, so any draw can be computed
directly from (seed, k).

Workloads (the reference's layered benchmark shape, benchmarks/time_*/):

* ``layered_circuit(n, layers, seed)``: per layer, H on every qubit, then RX,
  RY and RZ on every qubit with angles U[0, 2pi) drawn in insertion order,
  then a CNOT ladder i -> i+1.  n=20/10 layers is the 990-gate CPU workload;
  n=30/5 layers is the 745-gate, 16 GiB GPU workload.
* ``ansatz_circuit(n, layers, seed)``: per layer RY(theta) and RZ(theta) on
  every qubit, then a CNOT ring i -> (i+1) mod n, with the Z/ZZ chain
  observable H = sum a_i Z_i + sum_{i<n-1} b_i Z_i Z_{i+1}, a, b ~ N(0, 1).
* ``batch_workload(count, seed)``: ``count`` ansatz circuits with n uniform in
  [12, 18] and layer count uniform in [4, 12].
* ``stress_circuit(n, layers, seed)``: per layer U3 on every qubit, then 16
  CNOT or CZ gates on seeded distinct random pairs.
* ``rx_layers_circuit(n, layers)``: the reference examples' workload, L layers
  of RX(pi/(i+1)) on qubit i (examples/single_node_cpu/single_node.py:6-12),
  whose output is a product state with a closed form.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15

GATE_H, GATE_X, GATE_Y, GATE_Z, GATE_RX, GATE_RY, GATE_RZ = 0, 1, 2, 3, 4, 5, 6
GATE_CNOT, GATE_S, GATE_T, GATE_ID, GATE_U3, GATE_CZ = 7, 8, 9, 10, 11, 12

PARAMETRIC = (GATE_RX, GATE_RY, GATE_RZ)


def mix64(z: int) -> int:
    z &= MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


class SplitMix64:
    """Sequential splitmix64 stream."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + GOLDEN) & MASK64
        return mix64(self.state)

    def uniform(self) -> float:
        """U[0, 1) with 53 random bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def uniform_open0(self) -> float:
        """U(0, 1] with 53 random bits."""
        return ((self.next() >> 11) + 1) * (1.0 / (1 << 53))

    def normal(self) -> float:
        """N(0, 1) by Box-Muller (one value per two draws)."""
        u1 = self.uniform_open0()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi] (modulo reduction)."""
        return lo + self.next() % (hi - lo + 1)


def sample_uniforms(seed: int, k: int) -> np.ndarray:
    """The k uniforms in (0, 1] that the device sampler draws for `seed`.

    u_s = ((mix64(seed + (s+1)*GOLDEN) >> 11) + 1) * 2^-53, vectorised.
    """
    s = np.arange(1, k + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + s * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))


def layered_circuit(n: int, layers: int, seed: int):
    rng = SplitMix64(seed)
    gates = []
    for _ in range(layers):
        for q in range(n):
            gates.append((GATE_H, -1, q, ()))
        for kind in (GATE_RX, GATE_RY, GATE_RZ):
            for q in range(n):
                gates.append((kind, -1, q, (rng.uniform() * 2.0 * math.pi,)))
        for q in range(n - 1):
            gates.append((GATE_CNOT, q, q + 1, ()))
    return gates


def ansatz_circuit(n: int, layers: int, seed: int):
    """Returns (gates, a, b) with the chain observable coefficients."""
    rng = SplitMix64(seed)
    gates = []
    for _ in range(layers):
        for kind in (GATE_RY, GATE_RZ):
            for q in range(n):
                gates.append((kind, -1, q, (rng.uniform() * 2.0 * math.pi,)))
        for q in range(n):
            gates.append((GATE_CNOT, q, (q + 1) % n, ()))
    a = np.array([rng.normal() for _ in range(n)])
    b = np.array([rng.normal() for _ in range(n - 1)])
    return gates, a, b


def chain_pairs(n: int):
    return [(i, i + 1) for i in range(n - 1)]


def batch_workload(count: int, seed: int, n_range=(12, 18), layer_range=(4, 12)):
    """[(n, gates, a, b), ...] for `count` independent ansatz circuits."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = rng.randint(*n_range)
        layers = rng.randint(*layer_range)
        sub_seed = rng.next()
        gates, a, b = ansatz_circuit(n, layers, sub_seed)
        out.append((n, gates, a, b))
    return out


def stress_circuit(n: int, layers: int, seed: int, pairs_per_layer: int = 16):
    rng = SplitMix64(seed)
    gates = []
    for _ in range(layers):
        for q in range(n):
            th = rng.uniform() * 2.0 * math.pi
            ph = rng.uniform() * 2.0 * math.pi
            la = rng.uniform() * 2.0 * math.pi
            gates.append((GATE_U3, -1, q, (th, ph, la)))
        for _ in range(pairs_per_layer):
            c = rng.randint(0, n - 1)
            t = rng.randint(0, n - 2)
            if t >= c:
                t += 1
            kind = GATE_CNOT if rng.next() & 1 else GATE_CZ
            gates.append((kind, c, t, ()))
    return gates


def rx_layers_circuit(n: int, layers: int):
    gates = []
    for _ in range(layers):
        for i in range(n):
            gates.append((GATE_RX, -1, i, (math.pi / (i + 1),)))
    return gates


def rx_layers_closed_form(n: int, layers: int, indices) -> np.ndarray:
    """Amplitudes of rx_layers_circuit at the given basis indices.

    L layers of RX(pi/(i+1)) on qubit i equal RX(L*pi/(i+1)); the product
    state has amplitude prod_i (cos(t_i/2) if bit_i = 0 else -i sin(t_i/2)).
    """
    idx = np.asarray(indices, dtype=np.int64)
    out = np.ones(idx.shape, dtype=np.complex128)
    for i in range(n):
        t = layers * math.pi / (i + 1)
        bit = (idx >> i) & 1
        out *= np.where(bit == 0, math.cos(t / 2), -1j * math.sin(t / 2))
    return out


def count_parameters(gates) -> int:
    return sum(1 for g in gates if g[0] in PARAMETRIC)


def algorithmic_bytes_per_gate(gate, n: int, amp_bytes: int = 16) -> int:
    """2*S*N_touched for a one-gate HBM pass (uncontrolled: N, one control: N/2)."""
    kind, control, _target, _ = gate
    n_amp = 1 << n
    if kind == GATE_CZ:
        return 2 * amp_bytes * (n_amp // 4)
    if control >= 0:
        return 2 * amp_bytes * (n_amp // 2)
    return 2 * amp_bytes * n_amp
