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
