"""The five benchmark workloads at their full sizes on the GPU, through the
C ABI, against the oracle (oracle/oracle.c, the reference's multithreading
path circuit_multithreading.rs:9-54) or, where a 2^33 CPU state cannot
exist, against size-independent properties.

* 20 qubits, 990 gates (the CPU workload): amplitudes, probabilities and
  <Z_i> against the oracle run here, 1e-12; samples on injected uniforms
  equal to the oracle's inverse-CDF sampler (src/utils.rs:258-277).
* 30 qubits, 745 gates (16 GiB): <Z_i>, probabilities at 4096 seeded
  indices, the 10-qubit marginals of the top and the low qubits, and
  amplitudes at the seeded indices against tests/golden/config2_oracle.npz
  (written by tests/golden/make_config2.py), 1e-12; fused and per-gate.
* 24 qubits, 960 parameters: adjoint energy and <Z_i> against the oracle
  fixture tests/golden/config3_oracle.npz (1e-11 / 1e-12), the gradient
  against the oracle's parameter shift on 16 seeded parameters (1e-10) and
  against central finite differences of the GPU energy (h = 1e-6).
* 4096 circuits of 12-18 qubits in one launch: two circuits of every qubit
  count against oracle energies (1e-10) and two angles each by oracle
  parameter shift (1e-10) in complex128; complex64 against the same FP64
  oracle values, with the error printed and bounded.
* 33 qubits (128 GiB): norm after U, U^dagger U back to |0> (<Z_i> = 1),
  10^5 samples in range and consistent with <Z_i>, then the RX-layer
  closed form at 33 qubits.
"""
import ctypes
import gc
import math
import os
import sys

import numpy as np
import pytest

import oracle
from damavand_amd import _native as N
from damavand_amd import workloads as wl
from damavand_amd.circuit import Circuit, _CircuitHandle
from tests.libhelpers import amplitudes_at

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(n, gates, dtype="complex128"):
    c = Circuit(n, apply_method="gpu", dtype=dtype)
    for g in gates:
        c.add_gate(*g)
    return c


def inverse_gates(gates):
    inv = []
    for kind, c, t, p in reversed(gates):
        if kind in (wl.GATE_RX, wl.GATE_RY, wl.GATE_RZ):
            inv.append((kind, c, t, (-p[0],)))
        elif kind == wl.GATE_U3:
            th, ph, la = p
            inv.append((kind, c, t, (-th, -la, -ph)))
        elif kind in (wl.GATE_H, wl.GATE_X, wl.GATE_Y, wl.GATE_Z, wl.GATE_CNOT, wl.GATE_CZ):
            inv.append((kind, c, t, p))
        else:
            raise ValueError(kind)
    return inv


def shift_gradient_entries(n, gates, a, pairs, b, which):
    """Oracle parameter-shift gradient of the parameters `which` (indices in
    insertion order): one oracle forward keeping the state in front of each
    chosen gate, then only the gates after it per shifted energy."""
    pgate = [i for i, g in enumerate(gates) if g[0] in wl.PARAMETRIC]
    want = {pgate[k]: k for k in which}
    st = oracle.OracleState(n)
    saved = {}
    for gi, (kind, ctl, tgt, prm) in enumerate(gates):
        if gi in want:
            saved[gi] = st.amp.copy()
        st.apply_gate(kind, tgt, ctl, prm)
    energy = st.expect_zz(a, pairs, b)
    out = {}
    for k in which:
        gi = pgate[k]
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            st.amp[:] = saved[gi]
            kind, ctl, tgt, prm = gates[gi]
            st.apply_gate(kind, tgt, ctl, (prm[0] + s,))
            for kind2, ctl2, tgt2, prm2 in gates[gi + 1:]:
                st.apply_gate(kind2, tgt2, ctl2, prm2)
            vals.append(st.expect_zz(a, pairs, b))
        out[k] = 0.5 * (vals[0] - vals[1])
    return energy, out


# ------------------------------------------------------------- 20 qubits --
def test_config1_20q_layered(native):
    n, layers = 20, 10
    say(f"case 20q layered x{layers}")
    gates = wl.layered_circuit(n, layers, 1001)
    assert len(gates) == 990
    st = oracle.run_gates(n, gates)
    for per_gate in (False, True):
        c = build(n, gates)
        c.forward(per_gate=per_gate)
        amp = c.retrieve_amplitudes_on_host()
        ea = np.max(np.abs(amp - st.amp))
        ep = np.max(np.abs(c.measure() - np.abs(st.amp) ** 2))
        ez = np.max(np.abs(c.expectation_z() - st.expect_z()))
        say(f"  per_gate={per_gate}: amplitudes {ea:.3e} probabilities {ep:.3e} <Z> {ez:.3e}")
        assert ea <= 1e-12 and ep <= 1e-12 and ez <= 1e-12
    u = wl.sample_uniforms(77, 1000)
    got = c.state.sample_uniforms(u)
    ref = oracle.sample_uniforms(np.abs(st.amp) ** 2, u)
    mism = np.flatnonzero(got != ref)
    if mism.size:  # only a draw within 1e-12 of a cumulative boundary may differ
        cum = np.cumsum(np.abs(st.amp) ** 2)
        for j in mism:
            x = u[j] * cum[-1]
            assert min(abs(x - cum[got[j]]), abs(x - cum[ref[j]])) <= 1e-12, j
    say(f"  1000 samples on injected uniforms: {mism.size} boundary differences")


# ------------------------------------------------------------- 30 qubits --
@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "config2_oracle.npz")),
                    reason="tests/golden/config2_oracle.npz not generated")
def test_config2_30q_layered_vs_oracle(native):
    ref = np.load(os.path.join(GOLDEN, "config2_oracle.npz"))
    n, layers, seed = int(ref["n"]), int(ref["layers"]), int(ref["seed"])
    gates = wl.layered_circuit(n, layers, seed)
    assert len(gates) == int(ref["gates"]) == 745
    c = build(n, gates)
    idx = ref["idx"]
    for per_gate in (False, True):
        say(f"case 30q layered per_gate={per_gate}")
        c.reset_amplitudes()
        c.forward(per_gate=per_gate)
        ez = np.max(np.abs(c.expectation_z() - ref["expect_z"]))
        amp = amplitudes_at(c.state.handle, n, idx)
        ea = np.max(np.abs(amp - ref["amp"]))
        say(f"  <Z> {ez:.3e}, amplitudes at {idx.size} indices {ea:.3e}")
        assert ez <= 1e-12 and ea <= 1e-12
        if per_gate:
            continue
        p = c.measure()  # 8 GiB of probabilities on the host
        ep = np.max(np.abs(p[idx] - ref["prob"]))
        lo = p.reshape(-1, 1024).sum(axis=0)
        hi = p.reshape(1024, -1).sum(axis=1)
        em = max(np.max(np.abs(lo - ref["marg_lo"])), np.max(np.abs(hi - ref["marg_hi"])))
        en = abs(float(hi.sum()) - float(ref["norm"]))
        del p
        say(f"  probabilities at indices {ep:.3e}, 10-qubit marginals {em:.3e}, norm {en:.3e}")
        assert ep <= 1e-12 and em <= 1e-12 and en <= 1e-12


# ------------------------------------------------------------- adjoint --
@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "config3_oracle.npz")),
                    reason="tests/golden/config3_oracle.npz not generated")
def test_config3_24q_adjoint_vs_oracle_and_fd(native):
    ref = np.load(os.path.join(GOLDEN, "config3_oracle.npz"))
    n, layers, seed = int(ref["n"]), int(ref["layers"]), int(ref["seed"])
    gates, a, b = wl.ansatz_circuit(n, layers, seed)
    pairs = wl.chain_pairs(n)
    assert len(gates) == int(ref["gates"]) == 1440
    say(f"case adjoint {n}q x{layers}: 960 parameters")
    c = build(n, gates)
    e, g = c.gradient(a, pairs, b)
    assert g.size == 960
    ee = abs(e - float(ref["energy"]))
    idx = ref["idx"]
    eg = np.max(np.abs(g[idx] - ref["grad_shift"]))
    c.forward()
    ez = np.max(np.abs(c.expectation_z() - ref["expect_z"]))
    say(f"  energy {ee:.3e}, <Z> {ez:.3e}, gradient vs oracle parameter shift {eg:.3e} on {idx.size} parameters")
    assert ee <= 1e-11 and ez <= 1e-12 and eg <= 1e-10
    # central finite differences of the GPU energy, h = 1e-6, on a seeded subset
    h = 1e-6
    theta = c.get_parameters()
    rng = wl.SplitMix64(seed + 13)
    fd_idx = sorted({rng.next() % 960 for _ in range(8)})
    worst = 0.0
    for k in fd_idx:
        vals = []
        for s in (h, -h):
            t2 = theta.copy()
            t2[k] += s
            c.set_parameters(t2)
            c.reset_amplitudes()
            c.forward()
            vals.append(c.expectation(a, pairs, b))
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(fd - g[k]))
    c.set_parameters(theta)
    say(f"  gradient vs central differences (h=1e-6) on {len(fd_idx)} parameters: {worst:.3e}")
    # truncation h^2 |E'''| / 6 ~ 1e-13; rounding ~ 1e-16 x sum|coef| / h ~ 1e-9
    assert worst <= 1e-7


# --------------------------------------------------------------- batch --
def test_config4_batch_4096_vs_oracle(native):
    from damavand_amd import batch_forward_adjoint
    circuits = wl.batch_workload(4096, 1004)
    picks = []
    for q in range(12, 19):
        picks += [i for i, cc in enumerate(circuits) if cc[0] == q][:2]
    assert len(picks) == 14
    say("case batch 4096 circuits, complex128 and complex64 (one launch each)")
    e128, g128 = batch_forward_adjoint(circuits, dtype="complex128")
    e64, g64 = batch_forward_adjoint(circuits, dtype="complex64")
    worst = {"e128": 0.0, "g128": 0.0, "e64": 0.0, "g64": 0.0}
    for i in picks:
        n, gates, a, b = circuits[i]
        npar = wl.count_parameters(gates)
        which = [0, npar - 1]
        e_ref, gs = shift_gradient_entries(n, gates, a, wl.chain_pairs(n), b, which)
        worst["e128"] = max(worst["e128"], abs(e128[i] - e_ref))
        worst["e64"] = max(worst["e64"], abs(e64[i] - e_ref))
        for k in which:
            worst["g128"] = max(worst["g128"], abs(g128[i][k] - gs[k]))
            worst["g64"] = max(worst["g64"], abs(g64[i][k] - gs[k]))
        say(f"  circuit {i}: n={n}, {len(gates)} gates, energy c128 {abs(e128[i] - e_ref):.2e} "
            f"c64 {abs(e64[i] - e_ref):.2e}")
    say("  max errors vs FP64 oracle: " + ", ".join(f"{k} {v:.3e}" for k, v in worst.items()))
    assert worst["e128"] <= 1e-10 and worst["g128"] <= 1e-10
    # complex64: FP32 amplitudes, FP64 accumulation of the reductions
    assert worst["e64"] <= 1e-4 and worst["g64"] <= 1e-4
    # every circuit, c64 against c128
    d_e = float(np.max(np.abs(e64 - e128)))
    d_g = max(float(np.max(np.abs(x - y))) for x, y in zip(g64, g128))
    say(f"  all 4096: complex64 vs complex128 energy {d_e:.3e} gradient {d_g:.3e}")
    assert d_e <= 1e-4 and d_g <= 1e-4


# -------------------------------------------------------------- 33 qubits --
def test_config5_33q_stress(native):
    n, layers = 33, 20
    gates = wl.stress_circuit(n, layers, 1005)
    assert len(gates) == 980
    say(f"case stress {n}q: U, samples, then U^dagger on the same 128 GiB state")
    c = build(n, gates)
    c.forward()
    norm = c.state.norm_sqr()
    ez = c.expectation_z()
    say(f"  after U: |norm-1| {abs(norm - 1):.3e}")
    assert abs(norm - 1.0) < 1e-11 and np.all(np.abs(ez) <= 1 + 1e-12)
    k = 100000
    s = c.sample(k, seed=2024)
    assert s.dtype == np.int64 and s.min() >= 0 and s.max() < (1 << n)
    # sample means of Z_i against the exact <Z_i>: within 5.5 standard errors
    bits = (s[:, None] >> np.arange(n)[None, :]) & 1
    zbar = 1.0 - 2.0 * bits.mean(axis=0)
    sd = np.sqrt(np.maximum(1.0 - ez ** 2, 1e-12) / k)
    worst = float(np.max(np.abs(zbar - ez) / sd))
    say(f"  {k} samples in range, {np.unique(s).size} distinct, worst |mean Z - <Z>| = {worst:.2f} sigma")
    assert worst < 5.5
    inv = _CircuitHandle(n)
    for kind, ctl, tgt, prm in inverse_gates(gates):
        p = np.zeros(3)
        p[: len(prm)] = prm
        N.call("dmv_circuit_add_gate", inv.h, kind, ctl, tgt, N.dptr(p), 0, None)
    N.call("dmv_circuit_forward", inv.h, c.state.handle, N.FORWARD_FUSED)
    err = float(np.max(np.abs(c.expectation_z() - 1.0)))
    say(f"  after U^dagger U: max |<Z_i> - 1| {err:.3e}")
    assert err < 1e-10
    c.state.close()
    del c, inv
    gc.collect()
    say(f"case RX layers closed form n={n}")
    rx = build(n, wl.rx_layers_circuit(n, 10))
    rx.forward()
    want = np.array([math.cos(10 * math.pi / (i + 1)) for i in range(n)])
    err = float(np.max(np.abs(rx.expectation_z() - want)))
    rng = wl.SplitMix64(33)
    idx = np.array(sorted({rng.next() >> 31 for _ in range(256)}), dtype=np.int64)
    ea = float(np.max(np.abs(amplitudes_at(rx.state.handle, n, idx) - wl.rx_layers_closed_form(n, 10, idx))))
    say(f"  <Z_i> {err:.3e}, amplitudes at {idx.size} indices {ea:.3e}")
    assert err < 1e-12 and ea < 1e-12
    rx.state.close()
