"""Helpers for tests that drive libdamavand-gpu through its C ABI."""
import ctypes
import os
import re

import numpy as np

from damavand_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "damavand_gpu.h")


def header_functions():
    """Names of the functions declared in include/damavand_gpu.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|double|char)\s*\*?\s*([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


class NativeCircuit:
    """A dmv_circuit built from a gate list, usable without a GPU."""

    def __init__(self, n, gates):
        self.n = n
        self.gates = list(gates)
        h = ctypes.c_void_p()
        N.call("dmv_circuit_create", n, ctypes.byref(h))
        self.h = h
        for kind, control, target, params in gates:
            p = np.zeros(3)
            p[: len(params)] = params
            N.call("dmv_circuit_add_gate", h, kind, control, target, N.dptr(p), 0, None)

    def __del__(self):
        N.lib().dmv_circuit_destroy(self.h)

    def plan_ops(self, dtype=N.C128, tile_qubits=0):
        cap = 1 << 16
        kind = np.zeros(cap, dtype=np.int32)
        target = np.zeros(cap, dtype=np.int32)
        ctrl = np.zeros(cap, dtype=np.uint64)
        m = np.zeros(8 * cap)
        pas = np.zeros(cap, dtype=np.int32)
        count = ctypes.c_int(0)
        N.call("dmv_circuit_plan_ops", self.h, dtype, tile_qubits, N.iptr(kind), N.iptr(target),
               ctrl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), N.dptr(m), N.iptr(pas), cap, ctypes.byref(count))
        c = count.value
        mats = m[: 8 * c].reshape(c, 4, 2)
        mats = mats[..., 0] + 1j * mats[..., 1]
        return kind[:c], target[:c], ctrl[:c].astype(np.int64), mats.reshape(c, 2, 2), pas[:c]

    def plan_stats(self, dtype=N.C128):
        p, ph, o = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        b = ctypes.c_double(0)
        N.call("dmv_circuit_plan_stats", self.h, dtype, ctypes.byref(p), ctypes.byref(ph), ctypes.byref(o),
               ctypes.byref(b))
        return p.value, ph.value, o.value, b.value


def hip_memcpy():
    """hipMemcpy of the HIP runtime the library itself is bound to (looked up
    through the library's own handle, which searches its dependencies: the
    system runtime, or torch's copy when torch was loaded first), so that
    raw copies see the library's allocations."""
    fn = N.lib().hipMemcpy
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    fn.restype = ctypes.c_int
    return fn


def amplitudes_at(state_handle, n, indices, amp_bytes=16):
    """Amplitudes of a device state at the given plain indices, read element
    by element from its buffer through the layout that dmv_state_layout
    reports (include/damavand_gpu.h), without copying the whole state."""
    from tests.tile_emulator import element_map
    ptr = ctypes.c_void_p()
    N.call("dmv_state_device_ptr", state_handle, ctypes.byref(ptr), None)
    sh, bb, fo = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    N.call("dmv_state_layout", state_handle, ctypes.byref(sh), ctypes.byref(bb), ctypes.byref(fo))
    idx = np.asarray(indices, dtype=np.int64)
    assert idx.min() >= 0 and idx.max() < (1 << n)
    el = element_map(idx, sh.value, bb.value, fo.value)
    memcpy = hip_memcpy()
    N.synchronize()
    dt = np.complex128 if amp_bytes == 16 else np.complex64
    out = np.empty(idx.size, dtype=dt)
    one = np.empty(1, dtype=dt)
    for j, e in enumerate(el):
        assert memcpy(one.ctypes.data, ptr.value + int(e) * amp_bytes, amp_bytes, 2) == 0
        out[j] = one[0]
    return out.astype(np.complex128)
