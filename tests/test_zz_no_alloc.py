"""dmv_sample_static (the sampler without device allocation) against
dmv_sample, through the C ABI: the same state, count and seed give the same
samples, on two calls in a row (the second smaller, so stale staged results
of the first call would show), plus the size limit."""
import ctypes

import numpy as np
import pytest

from damavand_amd import _native as N
from damavand_amd import workloads as wl

MAX_STATIC = 1 << 20


def _state_with_circuit(n, layers, seed, dtype=N.C128):
    h = ctypes.c_void_p()
    N.call("dmv_state_create", n, dtype, ctypes.byref(h))
    c = ctypes.c_void_p()
    N.call("dmv_circuit_create", n, ctypes.byref(c))
    for kind, control, target, params in wl.layered_circuit(n, layers, seed):
        p = np.zeros(3)
        p[: len(params)] = params
        N.call("dmv_circuit_add_gate", c, kind, control, target, N.dptr(p), 0, None)
    N.call("dmv_circuit_forward", c, h, N.FORWARD_FUSED)
    N.call("dmv_circuit_destroy", c)
    return h


def _sample(name, h, k, seed):
    out = np.full(k, -7, dtype=np.int64)
    N.call(name, h, k, seed, N.i64ptr(out))
    return out


@pytest.mark.gpu
def test_static_sampler_matches_allocating_sampler(native):
    print("[no-alloc] 18q c128 + 15q c64 states, k = 200000 then 3000", flush=True)
    h = _state_with_circuit(18, 2, 41)
    h2 = _state_with_circuit(15, 2, 42, dtype=N.C64)
    try:
        ref1 = _sample("dmv_sample", h, 200000, 99)
        got1 = _sample("dmv_sample_static", h, 200000, 99)
        assert np.array_equal(got1, ref1)
        assert got1.min() >= 0 and got1.max() < (1 << 18)
        # second call: smaller count, other state and seed
        ref2 = _sample("dmv_sample", h2, 3000, 7)
        got2 = _sample("dmv_sample_static", h2, 3000, 7)
        assert np.array_equal(got2, ref2)
        assert got2.min() >= 0 and got2.max() < (1 << 15)
        # the first state again, at the largest count allowed
        ref3 = _sample("dmv_sample", h, MAX_STATIC, 5)
        got3 = _sample("dmv_sample_static", h, MAX_STATIC, 5)
        assert np.array_equal(got3, ref3)
    finally:
        N.call("dmv_state_destroy", h)
        N.call("dmv_state_destroy", h2)


@pytest.mark.gpu
def test_static_sampler_rejects_oversize(native):
    h = _state_with_circuit(12, 1, 3)
    try:
        out = np.zeros(4, dtype=np.int64)
        rc = N.lib().dmv_sample_static(h, MAX_STATIC + 1, 1, N.i64ptr(out))
        assert rc == 1 and "at most" in N.last_error()
        rc = N.lib().dmv_sample_static(h, -1, 1, N.i64ptr(out))
        assert rc == 1
        assert np.array_equal(_sample("dmv_sample_static", h, 4, 3), _sample("dmv_sample", h, 4, 3))
    finally:
        N.call("dmv_state_destroy", h)
