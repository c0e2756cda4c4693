"""Different handles driven from different threads at once (the header's
promise): each thread builds its own circuit and states and repeats forward,
readback, <Z>, <H>, sampling and the adjoint gradient; every repetition must
equal the oracle (or the first repetition, for samples).  ctypes releases the
GIL during each library call, so the calls overlap."""
import math
import threading

import numpy as np
import pytest

import oracle
from damavand_amd import workloads as wl
from damavand_amd.circuit import Circuit

pytestmark = pytest.mark.gpu

REPS = 4


def build(n, gates):
    c = Circuit(n, apply_method="gpu")
    for g in gates:
        c.add_gate(*g)
    return c


def test_two_handles_from_two_threads(native):
    na, nb = 17, 15
    ga = wl.layered_circuit(na, 2, 501)
    gb, a, b = wl.ansatz_circuit(nb, 2, 502)
    pairs = wl.chain_pairs(nb)
    out = {"a": [], "b": [], "err": []}

    def work_a():
        try:
            c = build(na, ga)
            for _ in range(REPS):
                c.reset_amplitudes()
                c.forward()
                out["a"].append((c.retrieve_amplitudes_on_host(), c.expectation_z(), c.sample(5000, seed=9)))
        except Exception as e:  # reported by the main thread
            out["err"].append(repr(e))

    def work_b():
        try:
            c = build(nb, gb)
            for _ in range(REPS):
                e, g = c.gradient(a, pairs, b)
                c.reset_amplitudes()
                c.forward()
                out["b"].append((e, g, c.expectation(a, pairs, b), c.retrieve_amplitudes_on_host()))
        except Exception as e:
            out["err"].append(repr(e))

    ts = [threading.Thread(target=work_a), threading.Thread(target=work_b)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    assert not out["err"], out["err"]
    assert len(out["a"]) == REPS and len(out["b"]) == REPS
    ref_a = oracle.run_gates(na, ga)
    ref_b = oracle.run_gates(nb, gb)
    e_ref = ref_b.expect_zz(a, pairs, b)
    pidx = [i for i, x in enumerate(gb) if x[0] in wl.PARAMETRIC]
    g_ref = {}
    for j in (0, len(pidx) - 1):
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            g2 = list(gb)
            k, ctl, t, p = g2[pidx[j]]
            g2[pidx[j]] = (k, ctl, t, (p[0] + s,))
            vals.append(oracle.run_gates(nb, g2).expect_zz(a, pairs, b))
        g_ref[j] = 0.5 * (vals[0] - vals[1])
    for amp, ez, smp in out["a"]:
        assert np.max(np.abs(amp - ref_a.amp)) <= 1e-12
        assert np.max(np.abs(ez - ref_a.expect_z())) <= 1e-12
        assert np.array_equal(smp, out["a"][0][2])
    for e, g, e2, amp in out["b"]:
        assert abs(e - e_ref) <= 1e-11 and abs(e2 - e_ref) <= 1e-11
        assert all(abs(g[j] - v) <= 1e-10 for j, v in g_ref.items())
        assert np.max(np.abs(amp - ref_b.amp)) <= 1e-12
