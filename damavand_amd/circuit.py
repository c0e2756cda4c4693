"""Python surface of the statevector path: ``Circuit`` mirrors the PyO3 class
``damavand.Circuit`` of the reference (src/qubit_backend/circuit.rs:60-770)
for ``apply_method="gpu"``; ``StateVector`` is a device-resident state.

All computation runs in libdamavand-gpu.so (HIP kernels for gfx950) through
the C ABI of include/damavand_gpu.h.  There is no CPU path: without the
library or a GPU every call raises.

Deviations from the reference, all deliberate:

* Results come back as NumPy arrays (``measure``, ``sample``, state getters,
  ``extract_expectation_values``) instead of Python lists; they index, iterate
  and compare like the lists they replace and avoid building 2^30-element
  lists.
* The state stays on the device; ``forward`` does not copy it to the host
  (circuit.rs:373-374 does).  Getters copy on demand.
* ``measure``/``sample`` keep the state (the reference's GPU path destroys it
  with cudaDeviceReset, damavand-gpu/rust_communication.cu:246).
* ``sample`` takes an optional ``seed`` (the reference's thread_rng is unseeded).
* Errors raise Python exceptions (ValueError, MemoryError, DamavandError)
  where the reference panics.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _native as N

APPLY_METHODS = ("brute_force", "shuffle", "multithreading", "gpu", "distributed_cpu", "distributed_gpu")
DTYPES = {"complex128": N.C128, "c128": N.C128, "complex64": N.C64, "c64": N.C64}


def _dtype_code(dtype) -> int:
    if isinstance(dtype, int):
        return dtype
    key = str(dtype).replace("numpy.", "").replace("torch.", "")
    if key not in DTYPES:
        raise ValueError(f"dtype must be complex128 or complex64, got {dtype!r}")
    return DTYPES[key]


def expectation_signs(samples, targets) -> np.ndarray:
    """(len(samples), len(targets)) array: -1.0 where bit `target` of the
    sample is set, else +1.0 (the readout rule of circuit.rs:494-513)."""
    s = np.asarray(samples, dtype=np.int64).reshape(-1)
    q = np.asarray(list(targets), dtype=np.int64)
    if q.size == 0:
        return np.zeros((s.size, 0))
    bits = (s[:, None] >> q[None, :]) & 1
    return np.where(bits > 0, -1.0, 1.0)


def initialize_mpi() -> None:
    """Reference module function (src/lib.rs:16-19).  The single-GPU path has
    no MPI, so there is nothing to initialise."""
    return None


class StateVector:
    """A 2^n-amplitude state in device memory (interleaved complex)."""

    def __init__(self, num_qubits: int, dtype="complex128"):
        self.num_qubits = int(num_qubits)
        self.dtype = _dtype_code(dtype)
        h = ctypes.c_void_p()
        N.call("dmv_state_create", self.num_qubits, self.dtype, ctypes.byref(h))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.lib().dmv_state_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self) -> int:
        return 1 << self.num_qubits

    def reset(self, basis_index: int = 0) -> None:
        N.call("dmv_state_reset", self._h, int(basis_index))

    def amplitudes(self) -> np.ndarray:
        re = np.empty(self.size)
        im = np.empty(self.size)
        N.call("dmv_state_get", self._h, N.dptr(re), N.dptr(im))
        return re + 1j * im

    def real(self) -> np.ndarray:
        re = np.empty(self.size)
        N.call("dmv_state_get", self._h, N.dptr(re), None)
        return re

    def imag(self) -> np.ndarray:
        im = np.empty(self.size)
        N.call("dmv_state_get", self._h, None, N.dptr(im))
        return im

    def load(self, amplitudes) -> None:
        a = np.asarray(amplitudes, dtype=np.complex128).reshape(-1)
        if a.size != self.size:
            raise ValueError(f"expected {self.size} amplitudes, got {a.size}")
        re = np.ascontiguousarray(a.real)
        im = np.ascontiguousarray(a.imag)
        N.call("dmv_state_set", self._h, N.dptr(re), N.dptr(im))

    def copy_from(self, other: "StateVector") -> None:
        N.call("dmv_state_copy", self._h, other._h)

    def apply_matrix(self, m, target: int, controls: Iterable[int] = ()) -> None:
        m = np.asarray(m, dtype=np.complex128).reshape(4)
        re = np.ascontiguousarray(m.real)
        im = np.ascontiguousarray(m.imag)
        mask = 0
        for c in controls:
            mask |= 1 << int(c)
        N.call("dmv_apply_matrix", self._h, N.dptr(re), N.dptr(im), mask, int(target))

    def apply_gate(self, kind: int, target: int, control: int = -1, params=None) -> None:
        p = np.zeros(3) if params is None else np.ascontiguousarray(
            np.concatenate([np.atleast_1d(np.asarray(params, dtype=np.float64)), np.zeros(3)])[:3])
        N.call("dmv_apply_gate", self._h, int(kind), int(control), int(target), N.dptr(p))

    def probabilities(self) -> np.ndarray:
        p = np.empty(self.size)
        N.call("dmv_probabilities", self._h, N.dptr(p))
        return p

    def norm_sqr(self) -> float:
        v = ctypes.c_double(0)
        N.call("dmv_norm_sqr", self._h, ctypes.byref(v))
        return v.value

    def expect_z(self) -> np.ndarray:
        out = np.empty(self.num_qubits)
        N.call("dmv_expect_z", self._h, N.dptr(out))
        return out

    def expect_zz(self, a=None, pairs: Sequence = (), b=None) -> float:
        a_arr = None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        if a_arr is not None and a_arr.size != self.num_qubits:
            raise ValueError("a needs one coefficient per qubit")
        pi, pj, bb = _pairs_arrays(pairs, b)
        out = ctypes.c_double(0)
        N.call("dmv_expect_zz", self._h, None if a_arr is None else N.dptr(a_arr), len(pi), N.iptr(pi), N.iptr(pj),
               N.dptr(bb), ctypes.byref(out))
        return out.value

    def expect_diag(self, wires: Sequence[int], table) -> float:
        """Exact <H> for the diagonal H|x> = table[j(x)]|x>, j(x) the bits of x
        at ``wires`` (wires[0] the least significant), up to 12 wires."""
        w, t = _table_arrays(wires, table)
        out = ctypes.c_double(0)
        N.call("dmv_expect_diag", self._h, w.size, N.iptr(w), N.dptr(t), ctypes.byref(out))
        return out.value

    def inner(self, other: "StateVector") -> complex:
        out = np.zeros(2)
        N.call("dmv_inner_product", self._h, other._h, N.dptr(out))
        return complex(out[0], out[1])

    def sample(self, k: int, seed: int) -> np.ndarray:
        out = np.empty(int(k), dtype=np.int64)
        N.call("dmv_sample", self._h, int(k), int(seed) & ((1 << 64) - 1), N.i64ptr(out))
        return out

    def sample_uniforms(self, u) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty(u.size, dtype=np.int64)
        N.call("dmv_sample_uniforms", self._h, u.size, N.dptr(u), N.i64ptr(out))
        return out


def _table_arrays(wires, table):
    w = np.ascontiguousarray(wires, dtype=np.int32).reshape(-1)
    t = np.ascontiguousarray(table, dtype=np.float64).reshape(-1)
    if t.size != 1 << w.size:
        raise ValueError(f"a table over {w.size} wires has {1 << w.size} entries, got {t.size}")
    return w, t


def _pairs_arrays(pairs, b):
    """(pi, pj, b) arrays of the ZZ terms; one coefficient per pair."""
    pairs = list(pairs)
    pi = np.ascontiguousarray([int(p[0]) for p in pairs], dtype=np.int32)
    pj = np.ascontiguousarray([int(p[1]) for p in pairs], dtype=np.int32)
    bb = np.ascontiguousarray(b if pairs else [], dtype=np.float64).reshape(-1)
    if bb.size != len(pairs):
        raise ValueError("b needs one coefficient per pair")
    return pi, pj, bb


class _CircuitHandle:
    def __init__(self, n: int):
        h = ctypes.c_void_p()
        N.call("dmv_circuit_create", int(n), ctypes.byref(h))
        self.h = h

    def __del__(self):
        try:
            if self.h and self.h.value:
                N.lib().dmv_circuit_destroy(self.h)
        except Exception:
            pass


class Circuit:
    """Drop-in for ``damavand.Circuit(num_qubits, apply_method="gpu")``.

    Gates are appended with the ``add_*`` methods and run by ``forward``
    (circuit.rs:341-375) as fused HBM passes of hand-written HIP kernels.
    """

    def __init__(self, num_qubits: int, apply_method: Optional[str] = None, dtype="complex128"):
        method = "gpu" if apply_method is None else apply_method
        if method not in APPLY_METHODS:
            # circuit.rs:111-112 panics with this message
            raise ValueError(f"Apply method not recognized: {method}")
        if method != "gpu":
            raise NotImplementedError(
                f"apply_method={method!r} is one of the reference's CPU/MPI backends; this library "
                "implements apply_method='gpu' on one MI355X")
        self.num_qubits = int(num_qubits)
        if self.num_qubits < 1:
            raise ValueError("num_qubits must be at least 1")
        self.apply_method = method
        self.dtype = _dtype_code(dtype)
        # device-memory check (replaces the host-RAM check of circuit.rs:115-128)
        mq = ctypes.c_int(0)
        N.call("dmv_max_qubits", self.dtype, ctypes.byref(mq))
        if self.num_qubits > mq.value:
            raise MemoryError(f"System requires more memory to simulate {self.num_qubits} qubits. "
                              f"Currently limited to {mq.value} qubits.")
        self._c = _CircuitHandle(self.num_qubits)
        self.state = StateVector(self.num_qubits, self.dtype)
        self._scratch = None
        self.observables = []  # gate indices, as circuit.rs:72
        self._targets = []     # target qubit of every gate
        self.num_gpus_per_node = 1
        self.num_amplitudes_per_gpu = 1 << self.num_qubits

    # ------------------------------------------------------------- state --
    def reset_amplitudes(self) -> None:
        """circuit.rs:262-302: back to |0...0>."""
        self.state.reset(0)

    def reset(self) -> None:
        """circuit.rs:303-306: reset the state and remove every gate."""
        self.reset_amplitudes()
        N.call("dmv_circuit_clear", self._c.h)
        self.observables = []
        self._targets = []

    def retrieve_amplitudes_on_host(self) -> np.ndarray:
        """circuit.rs:380-405: the amplitudes as a complex array."""
        return self.state.amplitudes()

    def get_real_part_state(self) -> np.ndarray:
        return self.state.real()

    def get_imaginary_part_state(self) -> np.ndarray:
        return self.state.imag()

    # ------------------------------------------------------------- gates --
    def _add(self, kind: int, control: int, target: int, params=(), observable: bool = False) -> int:
        p = np.zeros(3)
        p[: len(params)] = params
        idx = ctypes.c_int(0)
        N.call("dmv_circuit_add_gate", self._c.h, int(kind), int(control), int(target), N.dptr(p),
               1 if observable else 0, ctypes.byref(idx))
        self._targets.append(int(target))
        if observable:
            self.observables.append(idx.value)
        return idx.value

    def add_hadamard_gate(self, active_qubit: int) -> None:
        self._add(N.GATE_H, -1, active_qubit)

    def add_rotation_x_gate(self, active_qubit: int, theta: float) -> None:
        self._add(N.GATE_RX, -1, active_qubit, (float(theta),))

    def add_rotation_y_gate(self, active_qubit: int, theta: float) -> None:
        self._add(N.GATE_RY, -1, active_qubit, (float(theta),))

    def add_rotation_z_gate(self, active_qubit: int, theta: float) -> None:
        self._add(N.GATE_RZ, -1, active_qubit, (float(theta),))

    def add_pauli_x_gate(self, active_qubit: int, is_observable: bool) -> None:
        self._add(N.GATE_X, -1, active_qubit, observable=is_observable)

    def add_pauli_y_gate(self, active_qubit: int, is_observable: bool) -> None:
        self._add(N.GATE_Y, -1, active_qubit, observable=is_observable)

    def add_pauli_z_gate(self, active_qubit: int, is_observable: bool) -> None:
        self._add(N.GATE_Z, -1, active_qubit, observable=is_observable)

    def add_cnot_gate(self, control_qubit: int, target_qubit: int) -> None:
        self._add(N.GATE_CNOT, control_qubit, target_qubit)

    # additions (no reference counterpart)
    def add_u3_gate(self, active_qubit: int, theta: float, phi: float, lam: float) -> None:
        self._add(N.GATE_U3, -1, active_qubit, (float(theta), float(phi), float(lam)))

    def add_cz_gate(self, control_qubit: int, target_qubit: int) -> None:
        self._add(N.GATE_CZ, control_qubit, target_qubit)

    def add_gate(self, kind: int, control: int, target: int, params=()) -> None:
        self._add(kind, control, target, tuple(params))

    def print_operations(self) -> None:
        """circuit.rs:695-699 prints the (never filled) `operations` list: nothing."""
        return None

    # -------------------------------------------------------- parameters --
    @property
    def num_parameters(self) -> int:
        c = ctypes.c_int(0)
        N.call("dmv_circuit_num_parameters", self._c.h, ctypes.byref(c))
        return c.value

    def set_parameters(self, parameters: Sequence[float]) -> None:
        """circuit.rs:308-322: the k-th value goes to the k-th RX/RY/RZ gate."""
        p = np.ascontiguousarray(np.asarray(parameters, dtype=np.float64).reshape(-1))
        N.call("dmv_circuit_set_parameters", self._c.h, N.dptr(p), p.size)

    def get_parameters(self) -> np.ndarray:
        p = np.empty(self.num_parameters)
        N.call("dmv_circuit_get_parameters", self._c.h, N.dptr(p), p.size)
        return p

    # ----------------------------------------------------------- running --
    def forward(self, per_gate: bool = False) -> None:
        """circuit.rs:341-375: apply the gates (observables skipped) to the state."""
        N.call("dmv_circuit_forward", self._c.h, self.state.handle,
               N.FORWARD_PER_GATE if per_gate else N.FORWARD_FUSED)

    def measure(self) -> np.ndarray:
        """circuit.rs:565-589: |a_i|^2 for every amplitude."""
        return self.state.probabilities()

    def sample(self, num_samples: Optional[int] = None, seed: Optional[int] = None) -> np.ndarray:
        """circuit.rs:434-458: num_samples (default 1000) basis indices drawn
        from the output distribution (inverse CDF, src/utils.rs:258-277)."""
        k = 1000 if num_samples is None else int(num_samples)
        if seed is None:
            seed = int.from_bytes(os.urandom(8), "little")
        return self.state.sample(k, seed)

    def extract_expectation_values(self, samples) -> np.ndarray:
        """circuit.rs:494-513: for each sample and each observable (insertion
        order), -1 if the observable's target bit is set, else +1.  As in the
        reference the Pauli type is not used (every observable is read in the
        Z basis)."""
        return expectation_signs(samples, [self._targets[i] for i in self.observables])

    def get_fidelity_between_two_states_with_parameters(self, parameters_1, parameters_2) -> float:
        """circuit.rs:753-769 with circuit_metrics.rs:12-33: |<psi(p1)|psi(p2)>|^2.
        Like the reference, it ends with reset(), which removes the gates."""
        other = StateVector(self.num_qubits, self.dtype)
        self.reset_amplitudes()
        self.set_parameters(parameters_1)
        self.forward()
        other.copy_from(self.state)
        self.reset_amplitudes()
        self.set_parameters(parameters_2)
        self.forward()
        ov = other.inner(self.state)
        other.close()
        self.reset()
        return abs(ov) ** 2

    # ------------------------------------------------- exact expectations --
    def expectation_z(self) -> np.ndarray:
        """Exact <Z_q> of the current state for every qubit."""
        return self.state.expect_z()

    def expectation(self, a=None, pairs: Sequence = (), b=None) -> float:
        """Exact <H>, H = sum_q a_q Z_q + sum_t b_t Z_i Z_j over `pairs`."""
        return self.state.expect_zz(a, pairs, b)

    def gradient(self, a=None, pairs: Sequence = (), b=None):
        """Adjoint-method gradient of E(theta) = <0|U^dag H U|0> with respect
        to the RX/RY/RZ parameters (in set_parameters order).  Returns
        (E, dE/dtheta).  The circuit's own state is left untouched."""
        if self._scratch is None:
            self._scratch = (StateVector(self.num_qubits, self.dtype), StateVector(self.num_qubits, self.dtype))
        psi, lam = self._scratch
        a_arr = np.zeros(self.num_qubits) if a is None else np.ascontiguousarray(a, dtype=np.float64)
        if a_arr.size != self.num_qubits:
            raise ValueError("a needs one coefficient per qubit")
        pi, pj, bb = _pairs_arrays(pairs, b)
        g = np.empty(max(1, self.num_parameters))
        e = ctypes.c_double(0)
        N.call("dmv_circuit_adjoint", self._c.h, psi.handle, lam.handle, N.dptr(a_arr), len(pi), N.iptr(pi),
               N.iptr(pj), N.dptr(bb), ctypes.byref(e), N.dptr(g))
        return e.value, g[: self.num_parameters].copy()

    def gradient_diag(self, wires: Sequence[int], table):
        """Adjoint-method gradient of E(theta) = <0|U^dag H U|0> for the
        diagonal H of StateVector.expect_diag.  Returns (E, dE/dtheta)."""
        if self._scratch is None:
            self._scratch = (StateVector(self.num_qubits, self.dtype), StateVector(self.num_qubits, self.dtype))
        psi, lam = self._scratch
        w, t = _table_arrays(wires, table)
        g = np.empty(max(1, self.num_parameters))
        e = ctypes.c_double(0)
        N.call("dmv_circuit_adjoint_diag", self._c.h, psi.handle, lam.handle, w.size, N.iptr(w), N.dptr(t),
               ctypes.byref(e), N.dptr(g))
        return e.value, g[: self.num_parameters].copy()

    # ------------------------------------------------------------ tuning --
    def plan_stats(self):
        """(passes, phases, ops, algorithmic HBM bytes) of one fused forward."""
        p, ph, o = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        b = ctypes.c_double(0)
        N.call("dmv_circuit_plan_stats", self._c.h, self.dtype, ctypes.byref(p), ctypes.byref(ph), ctypes.byref(o),
               ctypes.byref(b))
        return p.value, ph.value, o.value, b.value

    def autotune(self, adjoint: bool = False) -> int:
        """Measures the fused-pass launch shapes of this circuit's forward (or,
        with adjoint=True, of its gradient's reverse sweep) and keeps the
        fastest for this circuit, for later circuits of the same size in this
        process, and in the file named by DMV_TUNE_CACHE.  The candidates run
        on scratch states; the circuit's state is left as it was.  Returns the
        chosen tile qubits."""
        k = ctypes.c_int(0)
        if adjoint:
            N.call("dmv_autotune_adjoint", self._c.h, self.dtype, ctypes.byref(k))
        else:
            N.call("dmv_autotune", self._c.h, self.state.handle, ctypes.byref(k))
        return k.value

    def tile_shape(self, adjoint: bool = False) -> dict:
        """The launch shape the next forward (or reverse-sweep) plan uses."""
        v = [ctypes.c_int(0) for _ in range(4)]
        N.call("dmv_circuit_tile_shape", self._c.h, self.dtype, 1 if adjoint else 0, *[ctypes.byref(x) for x in v])
        return {"tile_qubits": v[0].value, "threads_log2": v[1].value, "max_gen_per_pass": v[2].value,
                "nontemporal": v[3].value}

    @property
    def handle(self):
        return self._c.h
