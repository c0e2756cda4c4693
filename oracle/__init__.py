"""CPU oracle for the damavand statevector path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or as the timed
CPU baseline).  The product library (``damavand_amd``) never imports, links or
calls it.

Contents
--------
``oracle.c`` (built to ``liboracle.so`` by ``make -C oracle``) restates the
reference's ``multithreading`` backend in C:

===========================  ==============================================  =====================================
function                     reference it follows                            pinned by (tests/test_oracle_*.py)
===========================  ==============================================  =====================================
``gate_matrix``              src/qubit_backend/gates.rs:40-697                scipy expm(-i t P/2), textbook H/X/Y/Z/S/T,
                                                                             unitarity; U3 reduces to RY/RZ products
``apply_gate``               src/qubit_backend/circuit_multithreading.rs      KAT circuit.rs:811-836; np.kron brute
                             :9-54, circuit.rs:781-795                        force (brute.py) on random circuits;
                                                                             closed form of the RX-layer workload
                                                                             (examples/*, benchmarks/time_*); GHZ
``apply_gate_mask``          (generalisation, multi-control)                  np.kron brute force with projectors
``measure``                  src/qubit_backend/circuit.rs:579-581             closed forms (H^n uniform, GHZ), sum=1
``sample_uniforms``          src/utils.rs:258-277, circuit.rs:434-485         hand-computed known answers incl. zero-
                                                                             probability entries and boundaries;
                                                                             chi-square against exact probabilities
``expect_z`` / ``expect_zz`` (new; exact value of circuit.rs:494-513 means)   RX/RY product-state closed forms
                                                                             (<Z> = cos theta), brute force
===========================  ==============================================  =====================================

``brute.py`` is an independent dense ``np.kron`` simulator (the reference's
``brute_force`` method, circuit_brute_force.rs:9-12 with utils.rs:119-176)
used only to pin the C oracle on small circuits.

No function of the oracle is "parity unpinned".
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DMV_ORACLE_LIBRARY: another build of oracle.c (the sanitizer build of
# `make -C oracle san`)
_LIB_PATH = os.environ.get("DMV_ORACLE_LIBRARY") or os.path.join(_HERE, "liboracle.so")
_lib = None

GATE_H, GATE_X, GATE_Y, GATE_Z, GATE_RX, GATE_RY, GATE_RZ = 0, 1, 2, 3, 4, 5, 6
GATE_CNOT, GATE_S, GATE_T, GATE_ID, GATE_U3, GATE_CZ = 7, 8, 9, 10, 11, 12


def build() -> None:
    """Compile oracle.c into liboracle.so (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_gate_matrix.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_gate_matrix.restype = ctypes.c_int
        L.orc_init_zero.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.orc_apply_gate.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, dp, dp,
                                     ctypes.c_int, ctypes.c_int]
        L.orc_apply_gate_mask.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, dp,
                                          dp, ctypes.c_uint64, ctypes.c_int]
        L.orc_measure.argtypes = [ctypes.c_void_p, ctypes.c_int64, dp]
        L.orc_sample_uniforms.argtypes = [dp, ctypes.c_int64, dp, ctypes.c_int64, i64p]
        L.orc_sample_uniforms_rebuild.argtypes = [dp, ctypes.c_int64, dp, ctypes.c_int64, i64p]
        L.orc_expect_z.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.orc_expect_zz.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, ctypes.c_int, ip, ip, dp]
        L.orc_expect_zz.restype = ctypes.c_double
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gate_matrix(kind: int, params=None) -> np.ndarray:
    """2x2 complex matrix of a gate (gates.rs restated in oracle.c)."""
    p = np.zeros(3) if params is None else np.ascontiguousarray(np.atleast_1d(params), dtype=np.float64)
    if p.size < 3:
        p = np.concatenate([p, np.zeros(3 - p.size)])
    re = np.zeros(4)
    im = np.zeros(4)
    if lib().orc_gate_matrix(kind, _dp(p), _dp(re), _dp(im)) != 0:
        raise ValueError(f"unknown gate kind {kind}")
    return (re + 1j * im).reshape(2, 2)


class OracleState:
    """A complex128 state vector driven through the C oracle."""

    def __init__(self, n_qubits: int):
        self.n = n_qubits
        self.amp = np.zeros(1 << n_qubits, dtype=np.complex128)
        self._partner = np.empty_like(self.amp)
        lib().orc_init_zero(self.amp.ctypes.data, self.amp.size)

    def apply_matrix(self, m: np.ndarray, target: int, control: int = -1) -> None:
        m = np.asarray(m, dtype=np.complex128).reshape(4)
        re = np.ascontiguousarray(m.real)
        im = np.ascontiguousarray(m.imag)
        lib().orc_apply_gate(self.amp.ctypes.data, self._partner.ctypes.data, self.amp.size,
                             _dp(re), _dp(im), int(control), int(target))

    def apply_matrix_mask(self, m: np.ndarray, target: int, control_mask: int) -> None:
        m = np.asarray(m, dtype=np.complex128).reshape(4)
        re = np.ascontiguousarray(m.real)
        im = np.ascontiguousarray(m.imag)
        lib().orc_apply_gate_mask(self.amp.ctypes.data, self._partner.ctypes.data, self.amp.size,
                                  _dp(re), _dp(im), int(control_mask), int(target))

    def apply_gate(self, kind: int, target: int, control: int = -1, params=None) -> None:
        if kind == GATE_CZ:
            # CZ(c, t): Z on t controlled by c (symmetric).
            self.apply_matrix(gate_matrix(GATE_Z), target, control)
            return
        self.apply_matrix(gate_matrix(kind, params), target, control)

    def probabilities(self) -> np.ndarray:
        p = np.empty(self.amp.size)
        lib().orc_measure(self.amp.ctypes.data, self.amp.size, _dp(p))
        return p

    def expect_z(self) -> np.ndarray:
        out = np.empty(self.n)
        lib().orc_expect_z(self.amp.ctypes.data, self.n, _dp(out))
        return out

    def expect_zz(self, a, pairs, b) -> float:
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        pi = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
        pj = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        return lib().orc_expect_zz(self.amp.ctypes.data, self.n, _dp(a), len(pairs),
                                   pi.ctypes.data_as(ip), pj.ctypes.data_as(ip), _dp(b))


def sample_uniforms(prob: np.ndarray, u: np.ndarray, rebuild: bool = False) -> np.ndarray:
    """Reference inverse-CDF sampler (utils.rs:258-277) on injected uniforms."""
    prob = np.ascontiguousarray(prob, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(u.size, dtype=np.int64)
    fn = lib().orc_sample_uniforms_rebuild if rebuild else lib().orc_sample_uniforms
    fn(_dp(prob), prob.size, _dp(u), u.size, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out


def run_gates(n_qubits: int, gates) -> OracleState:
    """Apply a gate list [(kind, control, target, params), ...] from |0>."""
    st = OracleState(n_qubits)
    for kind, control, target, params in gates:
        st.apply_gate(kind, target, control, params)
    return st


def num_threads() -> int:
    return lib().orc_num_threads()


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))
