"""A PennyLane-style device and QNode over the GPU statevector path.

The reference is consumed from PennyLane through the external
``pennylane-damavand`` plugin as ``qml.device("damavand.qubit", wires=...,
apply_method="gpu")`` (README.md:92-107, docs/src/pennylane-plugin/README.md:
15-37).  PennyLane is not a dependency here; this module gives the same usage
with its own small operation and measurement vocabulary:

    from damavand_amd import qml
    dev = qml.device("damavand.qubit", wires=4, apply_method="gpu")

    @qml.qnode(dev)
    def circuit(theta):
        qml.RX(theta[0], wires=0)
        qml.CNOT(wires=[0, 1])
        return [qml.expval(qml.PauliZ(w)) for w in range(4)]

    theta = torch.tensor([0.3], requires_grad=True)
    loss = sum(circuit(theta))
    loss.backward()                     # adjoint kernels, one sweep per output

Operations: Hadamard, PauliX, PauliY, PauliZ, S, T, Identity, RX, RY, RZ,
Rot, U3, CNOT, CZ.  Measurements: expval and var of PauliX/PauliY/PauliZ,
their tensor products (``@``) and Hamiltonian sums; probs(wires); sample
(with shots); state().  Without shots, expval / var / probs are exact and
differentiable with torch: the vector-Jacobian product of every output is
one adjoint sweep with a diagonal observable (dmv_circuit_adjoint_diag).
With shots, results are estimated from samples drawn on the GPU and are not
differentiable.
"""
from __future__ import annotations

import itertools
import math
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from .circuit import Circuit

try:  # torch is optional: without it results are NumPy arrays
    import torch
except ImportError:  # pragma: no cover - torch is installed in this image
    torch = None

MAX_TABLE_WIRES = 12
_ACTIVE: List["_Tape"] = []


# ------------------------------------------------------------------ tape --
class _Tape:
    def __init__(self):
        self.ops: List[Operation] = []


class Operation:
    """A gate (or, once handed to a measurement, an observable factor)."""

    def __init__(self, name: str, wires, params=()):
        self.name = name
        self.wires = list(wires) if isinstance(wires, (list, tuple, range)) else [wires]
        self.params = list(params)
        if _ACTIVE:
            _ACTIVE[-1].ops.append(self)

    def __matmul__(self, other):
        return _as_observable(self) @ other

    def __rmul__(self, c):
        return float(c) * _as_observable(self)

    def __repr__(self):
        return f"{self.name}({', '.join(map(str, self.params + self.wires))})"


def _op(name, nparams, nwires):
    def make(*args, wires=None):
        params = list(args[:nparams])
        w = wires if wires is not None else (args[nparams] if len(args) > nparams else None)
        if w is None:
            raise TypeError(f"{name} needs wires")
        w = list(w) if isinstance(w, (list, tuple, range)) else [w]
        if len(w) != nwires or len(params) != nparams:
            raise TypeError(f"{name} takes {nparams} parameters and {nwires} wires")
        return Operation(name, w, params)
    make.__name__ = name
    return make


Hadamard = _op("Hadamard", 0, 1)
PauliX = _op("PauliX", 0, 1)
PauliY = _op("PauliY", 0, 1)
PauliZ = _op("PauliZ", 0, 1)
S = _op("S", 0, 1)
T = _op("T", 0, 1)
Identity = _op("Identity", 0, 1)
RX = _op("RX", 1, 1)
RY = _op("RY", 1, 1)
RZ = _op("RZ", 1, 1)
Rot = _op("Rot", 3, 1)    # RZ(omega) RY(theta) RZ(phi), phi applied first
U3 = _op("U3", 3, 1)      # e^{i(phi+lam)/2} RZ(phi) RY(theta) RZ(lam)
CNOT = _op("CNOT", 0, 2)
CZ = _op("CZ", 0, 2)


# ----------------------------------------------------------- observables --
class Observable:
    """sum_k coeffs[k] * prod_{(w, p) in terms[k]} P_w with P in {X, Y, Z}."""

    def __init__(self, terms):
        self.terms = terms  # list of (coeff, {wire: "X"|"Y"|"Z"})

    def __matmul__(self, other):
        other = _as_observable(other)
        out = []
        for (c1, p1), (c2, p2) in itertools.product(self.terms, other.terms):
            if set(p1) & set(p2):
                raise ValueError("a tensor product acts on distinct wires")
            out.append((c1 * c2, {**p1, **p2}))
        return Observable(out)

    def __add__(self, other):
        return Observable(self.terms + _as_observable(other).terms)

    def __rmul__(self, c):
        return Observable([(float(c) * k, p) for k, p in self.terms])

    __mul__ = __rmul__


def _as_observable(o) -> Observable:
    if isinstance(o, Observable):
        return o
    if isinstance(o, Operation):
        if _ACTIVE and o in _ACTIVE[-1].ops:
            _ACTIVE[-1].ops.remove(o)  # an observable, not a gate
        if o.name == "Identity":
            return Observable([(1.0, {})])
        if o.name in ("PauliX", "PauliY", "PauliZ"):
            return Observable([(1.0, {o.wires[0]: o.name[-1]})])
    raise TypeError(f"{o!r} is not a supported observable (PauliX/Y/Z, Identity, their products and sums)")


def Hamiltonian(coeffs: Sequence[float], observables: Sequence) -> Observable:
    out = Observable([])
    for c, o in zip(coeffs, observables):
        out = out + float(c) * _as_observable(o)
    return out


# ---------------------------------------------------------- measurements --
class Measurement:
    def __init__(self, kind: str, obs: Optional[Observable] = None, wires=None):
        self.kind = kind
        self.obs = obs
        self.wires = None if wires is None else (list(wires) if isinstance(wires, (list, tuple, range)) else [wires])


def expval(obs) -> Measurement:
    return Measurement("expval", _as_observable(obs))


def var(obs) -> Measurement:
    return Measurement("var", _as_observable(obs))


def probs(wires=None, op=None) -> Measurement:
    if op is not None:
        raise NotImplementedError("probs(op=...): measure in the Z basis, or rotate the wires first")
    return Measurement("probs", wires=wires)


def sample(obs=None, wires=None) -> Measurement:
    return Measurement("sample", None if obs is None else _as_observable(obs), wires)


def state() -> Measurement:
    return Measurement("state")


# ---------------------------------------------------------------- device --
class Device:
    """``device("damavand.qubit", wires, apply_method="gpu", shots=None)``."""

    name = "damavand.qubit"

    def __init__(self, wires, apply_method: str = "gpu", shots: Optional[int] = None, dtype="complex128",
                 seed: Optional[int] = None):
        if apply_method != "gpu":
            raise NotImplementedError(
                f"apply_method={apply_method!r} is one of the reference's CPU/MPI backends; this device runs "
                "apply_method='gpu'")
        self.wire_labels = list(range(wires)) if isinstance(wires, int) else list(wires)
        if len(set(self.wire_labels)) != len(self.wire_labels):
            raise ValueError("wire labels must be distinct")
        self.num_wires = len(self.wire_labels)
        self.shots = shots
        self.dtype = dtype
        self._seed = seed if seed is not None else 0x5EED
        self._calls = 0
        self._circuits: Dict[tuple, Circuit] = {}

    def wire(self, label) -> int:
        try:
            return self.wire_labels.index(label)
        except ValueError:
            raise ValueError(f"wire {label!r} is not on the device (wires {self.wire_labels})") from None

    def circuit_for(self, key, gates) -> Circuit:
        """A library Circuit with this gate structure (kept between calls, its
        angles updated with set_parameters)."""
        c = self._circuits.get(key)
        if c is None:
            c = Circuit(self.num_wires, apply_method="gpu", dtype=self.dtype)
            for g in gates:
                c.add_gate(*g)
            self._circuits[key] = c
        return c

    def next_seed(self) -> int:
        self._calls += 1
        return (self._seed * 0x9E3779B97F4A7C15 + self._calls) & ((1 << 64) - 1)


def device(name: str, wires, apply_method: str = "gpu", shots: Optional[int] = None, **kw) -> Device:
    if name != "damavand.qubit":
        raise ValueError(f"unknown device {name!r}; this library provides 'damavand.qubit'")
    return Device(wires, apply_method=apply_method, shots=shots, **kw)


# ------------------------------------------------------------- execution --
_FIXED = {"Hadamard": N.GATE_H, "PauliX": N.GATE_X, "PauliY": N.GATE_Y, "PauliZ": N.GATE_Z, "S": N.GATE_S,
          "T": N.GATE_T, "Identity": N.GATE_ID}
_ROT = {"RX": N.GATE_RX, "RY": N.GATE_RY, "RZ": N.GATE_RZ}
# gates that take a measured wire's basis to Z (PennyLane's diagonalizing gates)
_DIAG = {"Z": [], "X": [(N.GATE_H, -1)], "Y": [(N.GATE_Z, -1), (N.GATE_S, -1), (N.GATE_H, -1)]}


class _Program:
    """A recorded circuit lowered to library gates: fixed gates and rotations
    whose angles are slots of a flat parameter vector (so that Rot and U3 are
    differentiable through their RZ/RY/RZ factors)."""

    def __init__(self, dev: Device, ops: List[Operation]):
        self.dev = dev
        self.structure = []   # (kind, control, target, slot or -1)
        self.values = []      # the parameter objects of each slot (floats or tensors)
        self.phase = []       # global phases of U3 gates: (slot_phi, slot_lam)
        for op in ops:
            w = [dev.wire(x) for x in op.wires]
            if op.name in _FIXED:
                self.structure.append((_FIXED[op.name], -1, w[0], -1))
            elif op.name in _ROT:
                self._rot(_ROT[op.name], w[0], op.params[0])
            elif op.name == "Rot":
                phi, th, om = op.params
                self._rot(N.GATE_RZ, w[0], phi)
                self._rot(N.GATE_RY, w[0], th)
                self._rot(N.GATE_RZ, w[0], om)
            elif op.name == "U3":
                th, phi, lam = op.params
                s_lam = self._rot(N.GATE_RZ, w[0], lam)
                self._rot(N.GATE_RY, w[0], th)
                s_phi = self._rot(N.GATE_RZ, w[0], phi)
                self.phase.append((s_phi, s_lam))
            elif op.name in ("CNOT", "CZ"):
                if w[0] == w[1]:
                    raise ValueError(f"{op.name} needs two distinct wires")
                self.structure.append((N.GATE_CNOT if op.name == "CNOT" else N.GATE_CZ, w[0], w[1], -1))
            else:  # pragma: no cover - every constructor above is handled
                raise NotImplementedError(op.name)

    def _rot(self, kind, wire, value) -> int:
        self.values.append(value)
        self.structure.append((kind, -1, wire, len(self.values) - 1))
        return len(self.values) - 1

    def run(self, theta: np.ndarray, basis: Dict[int, str]) -> Circuit:
        """The circuit followed by the gates that rotate `basis` to Z, run
        from |0>; the Circuit is returned with its state."""
        extra = [(k, c, w, ()) for w, b in sorted(basis.items()) for k, c in _DIAG[b]]
        gates = [(k, c, t, (theta[s],) if s >= 0 else ()) for k, c, t, s in self.structure] + extra
        key = (tuple((k, c, t, s >= 0) for k, c, t, s in self.structure), tuple(extra))
        circ = self.dev.circuit_for(key, gates)
        circ.set_parameters(theta)
        circ.reset_amplitudes()
        circ.forward()
        return circ


def _parity_table(coeff_terms, wires):
    """Diagonal table over `wires` of sum_k c_k prod_{w in Z-set_k} Z_w."""
    t = np.zeros(1 << len(wires))
    idx = np.arange(t.size)
    for c, zs in coeff_terms:
        par = np.zeros(t.size, dtype=np.int64)
        for w in zs:
            par ^= (idx >> wires.index(w)) & 1
        t += c * (1 - 2 * par)
    return t


def _groups(obs: Observable):
    """Terms grouped by the basis they need: [(basis {wire: X/Y}, wires, table)].
    A group spans at most MAX_TABLE_WIRES wires."""
    by_basis: Dict[tuple, list] = {}
    for c, paulis in obs.terms:
        basis = tuple(sorted((w, p) for w, p in paulis.items() if p != "Z"))
        by_basis.setdefault(basis, []).append((c, sorted(paulis)))
    out = []
    for basis, terms in by_basis.items():
        chunk, wires = [], []
        for c, ws in terms:
            new = sorted(set(wires) | set(ws))
            if len(new) > MAX_TABLE_WIRES and chunk:
                out.append((dict(basis), wires, _parity_table(chunk, wires)))
                chunk, new = [], sorted(ws)
            chunk.append((c, ws))
            wires = new
        if len(wires) > MAX_TABLE_WIRES:
            raise ValueError(f"an observable term spans more than {MAX_TABLE_WIRES} wires")
        out.append((dict(basis), wires, _parity_table(chunk, wires) if wires else np.array([sum(c for c, _ in chunk)])))
    return out


class _Runner:
    """Evaluates a tape's measurements and their vector-Jacobian products."""

    def __init__(self, dev: Device, program: _Program, measurements: List[Measurement]):
        self.dev = dev
        self.prog = program
        self.meas = measurements

    def _wires(self, m: Measurement):
        return list(range(self.dev.num_wires)) if m.wires is None else [self.dev.wire(w) for w in m.wires]

    def _obs_groups(self, m: Measurement):
        mapped = Observable([(c, {self.dev.wire(w): p for w, p in ps.items()}) for c, ps in m.obs.terms])
        return _groups(mapped)

    def evaluate(self, theta: np.ndarray) -> List[np.ndarray]:
        if self.dev.shots is not None:
            return self._evaluate_shots(theta)
        out = []
        for m in self.meas:
            if m.kind in ("expval", "var"):
                e = e2 = 0.0
                for basis, wires, table in self._obs_groups(m):
                    circ = self.prog.run(theta, basis)
                    if not wires:
                        e += float(table[0])
                        continue
                    e += circ.state.expect_diag(wires, table)
                    if m.kind == "var":
                        if len(self._obs_groups(m)) != 1:
                            raise NotImplementedError("var() of a sum of terms in different bases")
                        e2 = circ.state.expect_diag(wires, table * table)
                out.append(np.array(e if m.kind == "expval" else e2 - e * e))
            elif m.kind == "probs":
                w = self._wires(m)
                p = self.prog.run(theta, {}).measure()
                out.append(_marginal(p, w, self.dev.num_wires))
            elif m.kind == "state":
                amp = self.prog.run(theta, {}).retrieve_amplitudes_on_host()
                for s_phi, s_lam in self.prog.phase:
                    amp = amp * np.exp(0.5j * (theta[s_phi] + theta[s_lam]))
                out.append(amp)
            else:
                raise ValueError("sample() needs a device with shots")
        return out

    def _evaluate_shots(self, theta: np.ndarray) -> List[np.ndarray]:
        shots = int(self.dev.shots)
        n = self.dev.num_wires
        out = []
        for m in self.meas:
            if m.kind == "state":
                raise ValueError("state() is not available with shots")
            if m.kind == "probs" or (m.kind == "sample" and m.obs is None):
                s = self.prog.run(theta, {}).sample(shots, seed=self.dev.next_seed())
                w = self._wires(m)
                bits = (s[:, None] >> np.array(w)[None, :]) & 1
                if m.kind == "sample":
                    out.append(bits)
                else:
                    idx = (bits << np.arange(len(w))[None, ::-1]).sum(axis=1)
                    out.append(np.bincount(idx, minlength=1 << len(w)) / shots)
                continue
            vals = np.zeros(shots)
            for basis, wires, table in self._obs_groups(m):
                if not wires:
                    vals += table[0]
                    continue
                s = self.prog.run(theta, basis).sample(shots, seed=self.dev.next_seed())
                j = np.zeros(shots, dtype=np.int64)
                for k, w in enumerate(wires):
                    j |= ((s >> w) & 1) << k
                vals += table[j]
            if m.kind == "sample":
                out.append(vals)
            elif m.kind == "expval":
                out.append(np.array(vals.mean()))
            else:
                out.append(np.array(vals.var()))
        return out

    def vjp(self, theta: np.ndarray, grads: List[Optional[np.ndarray]], values: List[np.ndarray]) -> np.ndarray:
        """sum over outputs of <upstream grad, d output / d theta>, one adjoint
        sweep per (output, basis group) with a diagonal observable."""
        total = np.zeros(len(theta))
        for m, g, val in zip(self.meas, grads, values):
            if g is None or not np.any(g):
                continue
            if m.kind in ("expval", "var"):
                for basis, wires, table in self._obs_groups(m):
                    if not wires:
                        continue
                    # d var = d<O^2> - 2 <O> d<O>: one diagonal observable
                    t = table if m.kind == "expval" else table * table - 2.0 * float(self._mean(m, theta)) * table
                    circ = self.prog.run(theta, basis)
                    total += float(g) * circ.gradient_diag(wires, t)[1]
            elif m.kind == "probs":
                w = self._wires(m)
                if len(w) > MAX_TABLE_WIRES:
                    raise NotImplementedError(f"gradients of probs over more than {MAX_TABLE_WIRES} wires")
                # probs are indexed with wires[0] as the most significant bit
                t = np.asarray(g, dtype=np.float64).reshape((2,) * len(w)).transpose(
                    tuple(range(len(w)))[::-1]).reshape(-1)
                circ = self.prog.run(theta, {})
                total += circ.gradient_diag(w, t)[1]
            else:
                raise ValueError(f"{m.kind}() is not differentiable")
        return total

    def _mean(self, m: Measurement, theta) -> float:
        e = 0.0
        for basis, wires, table in self._obs_groups(m):
            e += float(table[0]) if not wires else self.prog.run(theta, basis).state.expect_diag(wires, table)
        return e


def _marginal(p: np.ndarray, wires: List[int], n: int) -> np.ndarray:
    """Marginal probabilities over `wires`, wires[0] the most significant bit
    (PennyLane's ordering)."""
    t = p.reshape((2,) * n)                      # axis a is qubit n-1-a
    axes = [n - 1 - w for w in wires]
    keep = set(axes)
    t = t.sum(axis=tuple(a for a in range(n) if a not in keep))
    order = sorted(axes)                         # remaining axes in increasing axis order
    t = np.transpose(t, [order.index(a) for a in axes])
    return t.reshape(-1)


# ----------------------------------------------------------------- qnode --
if torch is not None:
    class _QNodeFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, runner, theta_np, *params):
            vals = runner.evaluate(theta_np)
            ctx.runner = runner
            ctx.theta = theta_np
            ctx.vals = vals
            return tuple(torch.as_tensor(v, dtype=torch.float64) if np.isrealobj(v) else torch.as_tensor(v)
                         for v in vals)

        @staticmethod
        def backward(ctx, *grads):
            g = [None if x is None else x.detach().cpu().numpy() for x in grads]
            total = ctx.runner.vjp(ctx.theta, g, ctx.vals)
            return (None, None) + tuple(torch.tensor(v, dtype=torch.float64) for v in total)


class QNode:
    """A function of gate parameters run on a Device (see the module docstring)."""

    def __init__(self, func, dev: Device):
        self.func = func
        self.dev = dev

    def __call__(self, *args, **kwargs):
        tape = _Tape()
        _ACTIVE.append(tape)
        try:
            ret = self.func(*args, **kwargs)
        finally:
            _ACTIVE.pop()
        single = isinstance(ret, Measurement)
        meas = [ret] if single else list(ret)
        if not all(isinstance(m, Measurement) for m in meas):
            raise TypeError("a qnode returns a measurement or a sequence of measurements")
        prog = _Program(self.dev, tape.ops)
        runner = _Runner(self.dev, prog, meas)
        theta = np.array([float(v.detach()) if torch is not None and isinstance(v, torch.Tensor) else float(v)
                          for v in prog.values])
        tensors = [v for v in prog.values if torch is not None and isinstance(v, torch.Tensor)]
        if tensors and any(t.requires_grad for t in tensors) and self.dev.shots is None:
            params = [v if isinstance(v, torch.Tensor) else torch.tensor(float(v), dtype=torch.float64)
                      for v in prog.values]
            out = _QNodeFunction.apply(runner, theta, *params)
        else:
            out = runner.evaluate(theta)
            if tensors:
                out = [torch.as_tensor(v) for v in out]
        return out[0] if single else tuple(out)


def qnode(dev: Device):
    def wrap(func):
        return QNode(func, dev)
    return wrap
