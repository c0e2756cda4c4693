"""ctypes binding of libdamavand-gpu.so (the C ABI of include/damavand_gpu.h).

The library is loaded from this package directory, or from the path in
DMV_LIBRARY (another build of the same library, e.g. the sanitizer build of
``python -m damavand_amd.build --sanitize``).  There is no fallback: every
call fails with an exception when the library is missing or when a dmv_*
call reports an error.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DMV_LIBRARY") or os.path.join(_HERE, "libdamavand-gpu.so")

DMV_OK = 0
C128, C64 = 0, 1
FORWARD_FUSED, FORWARD_PER_GATE = 0, 1
KC_TILE, KC_GATE, KC_SMALL, KC_REDUCE, KC_SAMPLE, KC_ADJOINT, KC_BATCH, KC_COPY = range(8)
KC_NAMES = ["tile", "gate", "small", "reduce", "sample", "adjoint", "batch", "copy"]

GATE_H, GATE_X, GATE_Y, GATE_Z, GATE_RX, GATE_RY, GATE_RZ = 0, 1, 2, 3, 4, 5, 6
GATE_CNOT, GATE_S, GATE_T, GATE_ID, GATE_U3, GATE_CZ = 7, 8, 9, 10, 11, 12


class DamavandError(RuntimeError):
    """A dmv_* call returned a nonzero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[dmv status {code}] {message}")
        self.code = code


class NoDeviceError(DamavandError):
    pass


class OutOfDeviceMemory(DamavandError):
    pass


_lib = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_c_u64_p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p

# name: (restype, argtypes); restype None means int status checked by `call`
_SIGNATURES = {
    "dmv_last_error": (ctypes.c_char_p, []),
    "dmv_version": (ctypes.c_char_p, []),
    "dmv_device_count": (ctypes.c_int, [_c_int_p]),
    "dmv_device_memory": (ctypes.c_int, [_c_u64_p, _c_u64_p]),
    "dmv_max_qubits": (ctypes.c_int, [ctypes.c_int, _c_int_p]),
    "dmv_state_bytes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_u64_p]),
    "dmv_state_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmv_state_destroy": (ctypes.c_int, [_vp]),
    "dmv_state_num_qubits": (ctypes.c_int, [_vp, _c_int_p]),
    "dmv_state_reset": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "dmv_state_get": (ctypes.c_int, [_vp, _c_dbl_p, _c_dbl_p]),
    "dmv_state_set": (ctypes.c_int, [_vp, _c_dbl_p, _c_dbl_p]),
    "dmv_state_copy": (ctypes.c_int, [_vp, _vp]),
    "dmv_state_device_ptr": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _c_u64_p]),
    "dmv_state_layout": (ctypes.c_int, [_vp, _c_int_p, _c_int_p, _c_int_p]),
    "dmv_state_pieces": (ctypes.c_int, [_vp, _c_int_p]),
    "dmv_apply_matrix": (ctypes.c_int, [_vp, _c_dbl_p, _c_dbl_p, ctypes.c_uint64, ctypes.c_int]),
    "dmv_apply_gate": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "dmv_probabilities": (ctypes.c_int, [_vp, _c_dbl_p]),
    "dmv_norm_sqr": (ctypes.c_int, [_vp, _c_dbl_p]),
    "dmv_expect_z": (ctypes.c_int, [_vp, _c_dbl_p]),
    "dmv_expect_zz": (ctypes.c_int, [_vp, _c_dbl_p, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p]),
    "dmv_expect_diag": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p]),
    "dmv_inner_product": (ctypes.c_int, [_vp, _vp, _c_dbl_p]),
    "dmv_sample": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_uint64, _c_i64_p]),
    "dmv_sample_uniforms": (ctypes.c_int, [_vp, ctypes.c_int64, _c_dbl_p, _c_i64_p]),
    "dmv_sample_static": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_uint64, _c_i64_p]),
    "dmv_circuit_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmv_circuit_destroy": (ctypes.c_int, [_vp]),
    "dmv_circuit_add_gate": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p, ctypes.c_int,
                                            _c_int_p]),
    "dmv_circuit_num_gates": (ctypes.c_int, [_vp, _c_int_p]),
    "dmv_circuit_num_parameters": (ctypes.c_int, [_vp, _c_int_p]),
    "dmv_circuit_set_parameters": (ctypes.c_int, [_vp, _c_dbl_p, ctypes.c_int]),
    "dmv_circuit_get_parameters": (ctypes.c_int, [_vp, _c_dbl_p, ctypes.c_int]),
    "dmv_circuit_observables": (ctypes.c_int, [_vp, _c_int_p, ctypes.c_int, _c_int_p]),
    "dmv_circuit_clear": (ctypes.c_int, [_vp]),
    "dmv_circuit_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "dmv_circuit_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _c_dbl_p, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p,
                                           _c_dbl_p, _c_dbl_p]),
    "dmv_circuit_adjoint_diag": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p,
                                                _c_dbl_p]),
    "dmv_circuit_plan_stats": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p]),
    "dmv_circuit_plan_ops": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_u64_p, _c_dbl_p,
                                            _c_int_p, ctypes.c_int, _c_int_p]),
    "dmv_circuit_plan_json": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_char_p, ctypes.c_uint64, _c_u64_p]),
    "dmv_batch_forward_adjoint": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                                 _c_int_p, _c_dbl_p, _c_int_p, _c_dbl_p, _c_int_p, _c_dbl_p,
                                                 _c_dbl_p]),
    "dmv_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "dmv_profile_reset": (ctypes.c_int, []),
    "dmv_profile_get": (ctypes.c_int, [ctypes.c_int, _c_i64_p, _c_dbl_p, _c_dbl_p]),
    "dmv_copy_benchmark": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, _c_dbl_p]),
    "dmv_synchronize": (ctypes.c_int, []),
    "dmv_tune_set": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64]),
    "dmv_tune_get": (ctypes.c_int, [ctypes.c_char_p, _c_i64_p]),
    "dmv_autotune": (ctypes.c_int, [_vp, _vp, _c_int_p]),
    "dmv_autotune_adjoint": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p]),
    "dmv_tune_forget": (ctypes.c_int, []),
    "dmv_batch_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                          _c_int_p, _c_dbl_p, _c_int_p, _c_dbl_p, _c_int_p, _c_int_p]),
    "dmv_circuit_tile_shape": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p,
                                              _c_int_p]),
    # compatibility layer (reference symbol names)
    "get_number_of_available_gpus": (ctypes.c_int, []),
    "get_memory_for_gpu": (ctypes.c_double, [ctypes.c_int]),
    "peer_access_allowed": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "print_timers": (None, []),
    "init_quantum_state": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "apply_one_qubit_gate_gpu_local": (None, [_c_dbl_p, _c_dbl_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int]),
    "apply_one_qubit_gate_gpu_distributed": (None, [_c_dbl_p, _c_dbl_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int]),
    "measure_on_gpu": (None, [ctypes.c_int, _c_dbl_p]),
    "sequential_measure_on_gpu": (None, [ctypes.c_int, _c_dbl_p]),
    "concurrent_measure_on_gpu": (None, [ctypes.c_int, _c_dbl_p]),
    "retrieve_amplitudes_on_host": (None, [ctypes.c_int, _c_dbl_p, _c_dbl_p]),
    "load_amplitudes_local_on_device": (None, [ctypes.c_int, _c_dbl_p, _c_dbl_p]),
    "split_amplitudes_between_gpus": (None, [ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "exchange_amplitudes_between_gpus": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m damavand_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def last_error() -> str:
    return lib().dmv_last_error().decode(errors="replace")


def call(name: str, *args) -> None:
    rc = getattr(lib(), name)(*args)
    if rc != DMV_OK:
        msg = last_error()
        if rc == 3:
            raise NoDeviceError(rc, msg)
        if rc == 4:
            raise OutOfDeviceMemory(rc, msg)
        raise DamavandError(rc, msg)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_c_dbl_p)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_c_int_p)


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(_c_i64_p)


def device_count() -> int:
    c = ctypes.c_int(0)
    call("dmv_device_count", ctypes.byref(c))
    return c.value


def version() -> str:
    return lib().dmv_version().decode()


def tune_set(name: str, value: int) -> None:
    call("dmv_tune_set", name.encode(), int(value))


def tune_get(name: str) -> int:
    v = ctypes.c_int64(0)
    call("dmv_tune_get", name.encode(), ctypes.byref(v))
    return v.value


def profile_enable(on: bool = True) -> None:
    call("dmv_profile_enable", 1 if on else 0)


def profile_reset() -> None:
    call("dmv_profile_reset")


def profile_get(kc: int):
    """(launches, total_ms, algorithmic_bytes) of a kernel class."""
    n = ctypes.c_int64(0)
    ms = ctypes.c_double(0)
    b = ctypes.c_double(0)
    call("dmv_profile_get", kc, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b))
    return n.value, ms.value, b.value


def synchronize() -> None:
    call("dmv_synchronize")


def copy_benchmark(nbytes: int, repeats: int = 5) -> float:
    ms = ctypes.c_double(0)
    call("dmv_copy_benchmark", int(nbytes), int(repeats), ctypes.byref(ms))
    return ms.value


def device_memory():
    f = ctypes.c_uint64(0)
    t = ctypes.c_uint64(0)
    call("dmv_device_memory", ctypes.byref(f), ctypes.byref(t))
    return f.value, t.value
