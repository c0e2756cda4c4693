"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bars: amplitudes, probabilities and expectations within 1e-12 max-abs in
complex128; gradients within 1e-10 of the oracle's parameter-shift
gradient; samples exactly equal to the oracle's for the same uniforms (or
within 1e-12 of a cumulative boundary), and chi-square consistent for seeded
draws.  complex64 results are held to 1e-4 (reverse-sweep energies and
gradients to 1e-5 of the complex128 adjoint) and the error is printed.
"""
import ctypes
import math
import sys

import numpy as np
import pytest

import oracle
from damavand_amd import _native as N
from damavand_amd import workloads as wl
from damavand_amd.circuit import Circuit, StateVector
from tests.libhelpers import hip_memcpy
from tests.stat_helpers import chi2_pvalue, marginal_pvalues

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build_circuit(n, gates, dtype="complex128"):
    c = Circuit(n, apply_method="gpu", dtype=dtype)
    for kind, control, target, params in gates:
        c.add_gate(kind, control, target, params)
    return c


def random_gates(rng, n, count):
    kinds = [wl.GATE_H, wl.GATE_X, wl.GATE_Y, wl.GATE_Z, wl.GATE_RX, wl.GATE_RY, wl.GATE_RZ, wl.GATE_S, wl.GATE_T,
             wl.GATE_CNOT, wl.GATE_U3, wl.GATE_CZ]
    out = []
    for _ in range(count):
        k = kinds[rng.integers(len(kinds))]
        t = int(rng.integers(n))
        c = -1
        if k in (wl.GATE_CNOT, wl.GATE_CZ):
            if n == 1:
                k = wl.GATE_H
            else:
                c = int(rng.integers(n - 1))
                c += c >= t
        out.append((k, c, t, tuple(rng.uniform(0, 2 * math.pi, 3))))
    return out


def test_reference_known_answer(native):
    # circuit.rs:802-837: H(0), H(1), CNOT(0,1) -> all amplitudes 1/2
    c = Circuit(2, apply_method="gpu")
    c.add_hadamard_gate(0)
    c.add_hadamard_gate(1)
    c.add_cnot_gate(0, 1)
    c.forward()
    amp = c.retrieve_amplitudes_on_host()
    err = np.max(np.abs(amp - 0.5))
    say("KAT max-abs error", err)
    assert err <= 1e-15


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 12, 13, 14, 16, 17, 18, 20])
def test_random_circuit_fused(native, n):
    say(f"case random fused n={n}")
    rng = np.random.default_rng(100 + n)
    gates = random_gates(rng, n, 250)
    c = build_circuit(n, gates)
    c.forward()
    ref = oracle.run_gates(n, gates).amp
    err = np.max(np.abs(c.retrieve_amplitudes_on_host() - ref))
    say(f"  n={n} passes={c.plan_stats()[0]} max-abs error {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("n,layers", [(20, 6), (22, 6)])
def test_budget_chosen_per_circuit(native, n, layers):
    """Random U3 + CNOT/CZ layers: the forward plan takes GEN budget 12 by its
    pass and phase counts (every pass with split transposes), and matches the
    oracle; the same circuit with the budget fixed at 8 matches too."""
    say(f"case GEN budget per circuit n={n} layers={layers}")
    gates = wl.stress_circuit(n, layers, 1005)
    c = build_circuit(n, gates)
    shape = c.tile_shape()
    assert shape["max_gen_per_pass"] == 12, shape
    ref = oracle.run_gates(n, gates).amp
    c.forward()
    err = np.max(np.abs(c.retrieve_amplitudes_on_host() - ref))
    say(f"  budget 12: passes={c.plan_stats()[0]} max-abs error {err:.3e}")
    assert err <= TOL
    N.tune_set("max_gen_per_pass", 8)
    try:
        c8 = build_circuit(n, gates)
        c8.forward()
        err8 = np.max(np.abs(c8.retrieve_amplitudes_on_host() - ref))
        say(f"  budget 8: passes={c8.plan_stats()[0]} max-abs error {err8:.3e}")
    finally:
        N.tune_set("max_gen_per_pass", -1)
    assert err8 <= TOL


@pytest.mark.parametrize("n,grid,nt", [(18, 3, 0), (20, 16, 1), (21, 7, 0)])
def test_pipelined_tile_passes(native, n, grid, nt):
    """The persistent, software-pipelined tile kernel (forced onto a small
    grid so that each workgroup walks several tiles, odd counts included)."""
    say(f"case pipelined n={n} grid={grid} nontemporal={nt}")
    rng = np.random.default_rng(300 + n)
    gates = random_gates(rng, n, 250)
    c = build_circuit(n, gates)
    N.tune_set("tile_grid", grid)
    N.tune_set("nontemporal", nt)
    try:
        c.forward()
    finally:
        N.tune_set("tile_grid", -1)
        N.tune_set("nontemporal", -1)
    err = np.max(np.abs(c.retrieve_amplitudes_on_host() - oracle.run_gates(n, gates).amp))
    say(f"  max-abs error {err:.3e}")
    assert err <= TOL


TILE_CASES = [
    (20, 13, 9, 8, {}), (19, 12, 9, 4, {}), (20, 11, 8, 40, {}), (22, 12, 8, 2, {}), (20, 11, 6, 8, {}),
    (20, 10, 6, 16, {}), (21, 12, 8, 8, {"tile_xcd_log2": 0}), (20, 12, 8, 8, {"coalesce_qubits": 5}),
    (20, 12, 8, 8, {"register_renaming": 0}), (21, 11, 7, 8, {}), (21, 11, 7, 6, {"tile_xcd_log2": 0}),
    (20, 10, 8, 8, {}), (20, 12, 8, 8, {"half_lds": 0}),
]


@pytest.mark.parametrize("n,k,tlog,gen,extra", TILE_CASES)
def test_tile_shapes(native, n, k, tlog, gen, extra):
    """Every tile shape and launch option the tuner may pick (64- to
    512-thread workgroups, 10- to 13-qubit tiles, compute budgets per pass,
    XCD-contiguous tile order, wider coalescing runs, renaming off, whole-tile
    transposes in place of split ones)."""
    say(f"case tile shape n={n} k={k} tlog={tlog} gen={gen} {extra}")
    rng = np.random.default_rng(400 + n + k)
    gates = random_gates(rng, n, 300)
    c = build_circuit(n, gates)
    knobs = dict(tile_qubits=k, tile_threads_log2=tlog, max_gen_per_pass=gen, **extra)
    for key, v in knobs.items():
        N.tune_set(key, v)
    try:
        c.forward()
    finally:
        for key in knobs:
            N.tune_set(key, -1)  # unset: the default applies again
    err = np.max(np.abs(c.retrieve_amplitudes_on_host() - oracle.run_gates(n, gates).amp))
    say(f"  passes={c.plan_stats()[0]} max-abs error {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("n", [3, 12, 16])
def test_random_circuit_per_gate(native, n):
    say(f"case random per-gate n={n}")
    rng = np.random.default_rng(200 + n)
    gates = random_gates(rng, n, 150)
    c = build_circuit(n, gates)
    c.forward(per_gate=True)
    err = np.max(np.abs(c.retrieve_amplitudes_on_host() - oracle.run_gates(n, gates).amp))
    say(f"  n={n} max-abs error {err:.3e}")
    assert err <= TOL


def test_forward_does_not_reset_state(native):
    # circuit.rs:341-375 applies the gates on top of the current state
    gates = [(wl.GATE_RX, -1, 0, (0.3,)), (wl.GATE_CNOT, 0, 1, ())]
    c = build_circuit(2, gates)
    c.forward()
    c.forward()
    ref = oracle.run_gates(2, gates + gates).amp
    assert np.max(np.abs(c.retrieve_amplitudes_on_host() - ref)) <= TOL


def test_layered_workload_20q(native):
    """The 990-gate, 20-qubit layered circuit: amplitudes, probabilities, <Z_i>."""
    say("case layered 20q x 10 layers")
    n = 20
    gates = wl.layered_circuit(n, 10, 1001)
    c = build_circuit(n, gates)
    c.forward()
    st = oracle.run_gates(n, gates)
    e_amp = np.max(np.abs(c.retrieve_amplitudes_on_host() - st.amp))
    e_p = np.max(np.abs(c.measure() - st.probabilities()))
    e_z = np.max(np.abs(c.expectation_z() - st.expect_z()))
    say(f"  amplitudes {e_amp:.3e} probabilities {e_p:.3e} <Z> {e_z:.3e}")
    assert e_amp <= TOL and e_p <= TOL and e_z <= TOL
    # per-gate path on the same circuit
    c.reset_amplitudes()
    c.forward(per_gate=True)
    assert np.max(np.abs(c.retrieve_amplitudes_on_host() - st.amp)) <= TOL


def test_host_copies_in_pieces(native):
    """Probabilities and amplitudes of 22- and 21-qubit states (32 and 16 MiB
    per array, so four and two pieces through each state's pinned buffer,
    alternating its halves), read twice into fresh arrays, against the
    oracle; the two states keep their own buffers."""
    rng = np.random.default_rng(2201)
    runs = []
    for n in (22, 21):
        say(f"case host copies in pieces {n}q")
        gates = random_gates(rng, n, 60)
        c = build_circuit(n, gates)
        c.forward()
        runs.append((c, oracle.run_gates(n, gates)))
    for rep in range(2):
        for c, st in runs:
            e_p = np.max(np.abs(c.measure() - st.probabilities()))
            e_amp = np.max(np.abs(c.retrieve_amplitudes_on_host() - st.amp))
            say(f"  {c.num_qubits}q read {rep}: probabilities {e_p:.3e} amplitudes {e_amp:.3e}")
            assert e_p <= TOL and e_amp <= TOL


def test_observables_are_skipped(native):
    c = Circuit(3, apply_method="gpu")
    c.add_hadamard_gate(0)
    c.add_pauli_x_gate(1, True)
    c.add_pauli_z_gate(0, True)
    c.add_cnot_gate(0, 2)
    c.forward()
    ref = oracle.run_gates(3, [(wl.GATE_H, -1, 0, ()), (wl.GATE_CNOT, 0, 2, ())]).amp
    assert np.max(np.abs(c.retrieve_amplitudes_on_host() - ref)) <= TOL
    s = c.sample(2000, seed=5)
    ev = c.extract_expectation_values(s)
    assert ev.shape == (2000, 2)
    assert set(np.unique(s)) <= {0, 5}
    np.testing.assert_array_equal(ev[:, 0], 1.0)  # qubit 1 never set
    np.testing.assert_array_equal(ev[:, 1], np.where(s & 1, -1.0, 1.0))


def test_compat_layer_call_sequence(native):
    """The Rust crate's call sequence (circuit.rs:190-195, circuit_gpu.rs:42-58,
    circuit.rs:573-577, :387-391) against the oracle."""
    say("case compat layer")
    L = N.lib()
    n = 10
    n_amp = 1 << n
    assert L.get_number_of_available_gpus() >= 1
    assert L.get_memory_for_gpu(0) > 1000.0
    L.init_quantum_state(n_amp, 1, 1)
    rng = np.random.default_rng(9)
    gates = random_gates(rng, n, 120)
    for kind, control, target, params in gates:
        m = oracle.gate_matrix(wl.GATE_Z if kind == wl.GATE_CZ else kind, params).reshape(4)
        re = np.ascontiguousarray(m.real)
        im = np.ascontiguousarray(m.imag)
        L.apply_one_qubit_gate_gpu_local(N.dptr(re), N.dptr(im), n, n_amp, control, target)
    probs = np.zeros(n_amp)
    L.measure_on_gpu(n_amp, N.dptr(probs))
    re = np.zeros(n_amp)
    im = np.zeros(n_amp)
    L.retrieve_amplitudes_on_host(n_amp, N.dptr(re), N.dptr(im))
    st = oracle.run_gates(n, gates)
    assert np.max(np.abs(re + 1j * im - st.amp)) <= TOL
    assert np.max(np.abs(probs - st.probabilities())) <= TOL
    # the state survives measure (unlike rust_communication.cu:246); measure twice
    probs2 = np.zeros(n_amp)
    L.measure_on_gpu(n_amp, N.dptr(probs2))
    np.testing.assert_array_equal(probs, probs2)
    # re-init resets to |0>
    L.init_quantum_state(n_amp, 1, 1)
    L.retrieve_amplitudes_on_host(n_amp, N.dptr(re), N.dptr(im))
    assert re[0] == 1.0 and np.count_nonzero(re) == 1 and np.count_nonzero(im) == 0


def test_reductions(native):
    say("case reductions n=15")
    n = 15
    rng = np.random.default_rng(3)
    gates = random_gates(rng, n, 200)
    c = build_circuit(n, gates)
    c.forward()
    st = oracle.run_gates(n, gates)
    a = rng.normal(size=n)
    pairs = [(i, i + 1) for i in range(n - 1)] + [(0, 14), (3, 9)]
    b = rng.normal(size=len(pairs))
    e1 = abs(c.state.norm_sqr() - 1.0)
    e2 = np.max(np.abs(c.expectation_z() - st.expect_z()))
    e3 = abs(c.expectation(a, pairs, b) - st.expect_zz(a, pairs, b))
    say(f"  norm {e1:.3e} <Z> {e2:.3e} <H> {e3:.3e}")
    assert e1 <= 1e-12 and e2 <= TOL and e3 <= 1e-11


def test_sampler_matches_oracle_on_same_uniforms(native):
    say("case sampler injected uniforms")
    n = 13
    rng = np.random.default_rng(4)
    gates = random_gates(rng, n, 150)
    c = build_circuit(n, gates)
    c.forward()
    u = wl.sample_uniforms(2024, 20000)
    got = c.state.sample_uniforms(u)
    p = oracle.run_gates(n, gates).probabilities()
    ref = oracle.sample_uniforms(p, u)
    bad = np.nonzero(got != ref)[0]
    cum = np.concatenate([[0.0], np.cumsum(p)])
    for i in bad:  # only allowed within rounding of a cumulative boundary
        xsi = u[i] * cum[-1]
        assert np.min(np.abs(cum - xsi)) < 1e-12, (i, got[i], ref[i])
    say(f"  {len(bad)} of {u.size} draws differ, all at cumulative boundaries")
    # seeded draws use exactly the published uniform stream
    np.testing.assert_array_equal(c.state.sample(5000, 77), c.state.sample_uniforms(wl.sample_uniforms(77, 5000)))


def test_sampler_chi_square(native):
    say("case sampler chi-square")
    n = 10
    rng = np.random.default_rng(6)
    c = build_circuit(n, random_gates(rng, n, 100))
    c.forward()
    p = c.measure()
    s = c.sample(200000, seed=11)
    assert chi2_pvalue(s, p) > 1e-4
    # 20 qubits: top-bit buckets and per-qubit marginals
    n = 20
    c = build_circuit(n, wl.layered_circuit(n, 2, 5))
    c.forward()
    p = c.measure()
    s = c.sample(100000, seed=12)
    top = 10
    pb = p.reshape(1 << top, -1).sum(axis=1)
    assert chi2_pvalue(s >> (n - top), pb) > 1e-4
    assert min(marginal_pvalues(s, p, n)) > 1e-4 / n


def test_sampler_zero_state_raises(native):
    st = StateVector(4)
    st.load(np.zeros(16))
    with pytest.raises(N.DamavandError):
        st.sample(10, 1)


def _param_shift(n, gates, a, pairs, b):
    """Parameter-shift gradient on the oracle (exact for RX/RY/RZ)."""
    idx = [i for i, g in enumerate(gates) if g[0] in wl.PARAMETRIC]
    grad = np.zeros(len(idx))
    for k, gi in enumerate(idx):
        vals = []
        for s in (+0.5 * math.pi, -0.5 * math.pi):
            g2 = list(gates)
            kind, c, t, p = g2[gi]
            g2[gi] = (kind, c, t, (p[0] + s,))
            vals.append(oracle.run_gates(n, g2).expect_zz(a, pairs, b))
        grad[k] = 0.5 * (vals[0] - vals[1])
    return grad


@pytest.mark.parametrize("n,layers", [(5, 3), (11, 2), (13, 2), (16, 1)])
def test_adjoint_vs_parameter_shift(native, n, layers):
    say(f"case adjoint n={n} layers={layers}")
    gates, a, b = wl.ansatz_circuit(n, layers, 3000 + n)
    gates = gates + [(wl.GATE_RX, -1, n // 2, (0.37,))]
    pairs = wl.chain_pairs(n)
    c = build_circuit(n, gates)
    e, g = c.gradient(a, pairs, b)
    st = oracle.run_gates(n, gates)
    e_ref = st.expect_zz(a, pairs, b)
    g_ref = _param_shift(n, gates, a, pairs, b)
    ee, eg = abs(e - e_ref), np.max(np.abs(g - g_ref))
    say(f"  energy {ee:.3e} gradient max-abs {eg:.3e} over {g.size} parameters")
    assert ee <= 1e-11 and eg <= 1e-10
    # the circuit's own state is untouched by gradient()
    assert abs(c.state.norm_sqr() - 1.0) < 1e-14
    assert c.retrieve_amplitudes_on_host()[0] == 1.0


@pytest.mark.parametrize("n,layers,knobs", [
    (12, 2, {}), (18, 2, {}), (20, 2, {"adjoint_tile_qubits": 12}),
    (18, 2, {"adjoint_tile_qubits": 10, "tile_qubits": 10, "tile_threads_log2": 6})])
def test_adjoint_complex64(native, n, layers, knobs):
    """The reverse sweep in complex64 (FP32 amplitudes, FP64 sums) against
    the complex128 adjoint, which at 12 qubits is itself checked against the
    oracle's parameter shift.  The default tiles are 8 amplitudes per thread
    and state with 64 threads (12 qubits) or 256 threads (18); the tunables
    select 16 amplitudes per thread on 256 threads, and on 64."""
    say(f"case adjoint complex64 n={n} layers={layers} {knobs}")
    gates, a, b = wl.ansatz_circuit(n, layers, 3100 + n)
    pairs = wl.chain_pairs(n)
    e_ref, g_ref = build_circuit(n, gates).gradient(a, pairs, b)
    if n <= 12:
        assert np.max(np.abs(g_ref - _param_shift(n, gates, a, pairs, b))) <= 1e-10
    c64 = build_circuit(n, gates, dtype="complex64")
    for key, v in knobs.items():
        N.tune_set(key, v)
    try:
        shape = c64.tile_shape(adjoint=True)
        e, g = c64.gradient(a, pairs, b)
    finally:
        for key in knobs:
            N.tune_set(key, -1)
    ee, eg = abs(e - e_ref), float(np.max(np.abs(g - g_ref)))
    say(f"  complex64 vs complex128: energy {ee:.3e} gradient max-abs {eg:.3e} over {g.size} parameters "
        f"(largest |g| {np.max(np.abs(g_ref)):.3e})")
    assert g.shape == g_ref.shape
    if "adjoint_tile_qubits" in knobs:
        assert shape["tile_qubits"] == knobs["adjoint_tile_qubits"], shape
        assert shape["threads_log2"] == knobs.get("tile_threads_log2", 8), shape
    assert ee <= 1e-5 and eg <= 1e-5


def test_adjoint_narrow_tile(native):
    """The reverse sweep with 4 amplitudes per thread and state (k = 10, a
    tunable) against the default 8 (k = 11) on an 18-qubit ansatz."""
    say("case adjoint k=10 against k=11, n=18")
    n = 18
    gates, a, b = wl.ansatz_circuit(n, 3, 3300)
    pairs = wl.chain_pairs(n)
    e_ref, g_ref = build_circuit(n, gates).gradient(a, pairs, b)
    c = build_circuit(n, gates)
    N.tune_set("adjoint_tile_qubits", 10)
    try:
        e, g = c.gradient(a, pairs, b)
    finally:
        N.tune_set("adjoint_tile_qubits", -1)
    ee, eg = abs(e - e_ref), float(np.max(np.abs(g - g_ref)))
    say(f"  energy {ee:.3e} gradient max-abs {eg:.3e} over {g.size} parameters")
    assert ee <= 1e-12 and eg <= 1e-12


@pytest.mark.parametrize("dtype,knobs", [("complex128", {}), ("complex64", {}),
                                         ("complex64", {"adjoint_tile_qubits": 12})])
def test_adjoint_streaming(native, dtype, knobs):
    """Streaming (nontemporal) loads and stores in the reverse sweep change
    only the cache policy: energy and gradient are bit-identical to the
    cached form."""
    say(f"case adjoint streaming {dtype} {knobs}")
    n = 20
    gates, a, b = wl.ansatz_circuit(n, 2, 3400)
    pairs = wl.chain_pairs(n)
    out = []
    for nt in (0, 1):
        c = build_circuit(n, gates, dtype=dtype)
        for key, v in dict(knobs, nontemporal=nt).items():
            N.tune_set(key, v)
        try:
            assert c.tile_shape(adjoint=True)["nontemporal"] == nt
            out.append(c.gradient(a, pairs, b))
        finally:
            for key in dict(knobs, nontemporal=nt):
                N.tune_set(key, -1)
    (e0, g0), (e1, g1) = out
    say(f"  energy {e0!r} / {e1!r}, gradients equal: {np.array_equal(g0, g1)}")
    assert e0 == e1 and np.array_equal(g0, g1)


def test_adjoint_closed_form(native):
    # RY(t)|0>, H = Z: E = cos t, dE/dt = -sin t
    for t in (0.0, 0.4, 2.0, 4.5):
        c = Circuit(1, apply_method="gpu")
        c.add_rotation_y_gate(0, t)
        e, g = c.gradient(np.array([1.0]))
        assert abs(e - math.cos(t)) < 1e-15 and abs(g[0] + math.sin(t)) < 1e-15


def test_fidelity(native):
    n = 8
    gates, _, _ = wl.ansatz_circuit(n, 2, 41)
    c = build_circuit(n, gates)
    p1 = np.linspace(0, 3, c.num_parameters)
    p2 = p1 + 0.1
    f = c.get_fidelity_between_two_states_with_parameters(p1, p2)

    def run(p):
        g2, k = [], 0
        for kind, cc, t, pr in gates:
            if kind in wl.PARAMETRIC:
                pr = (p[k],)
                k += 1
            g2.append((kind, cc, t, pr))
        return oracle.run_gates(n, g2).amp

    ref = abs(np.vdot(run(p1), run(p2))) ** 2
    assert abs(f - ref) < 1e-12
    # circuit_metrics.rs:30: the gate list is cleared afterwards
    assert c.num_parameters == 0


def test_complex64_forward(native):
    say("case complex64")
    for n in (9, 14, 19):
        rng = np.random.default_rng(50 + n)
        gates = random_gates(rng, n, 200)
        c = build_circuit(n, gates, dtype="complex64")
        c.forward()
        err = np.max(np.abs(c.retrieve_amplitudes_on_host() - oracle.run_gates(n, gates).amp))
        say(f"  complex64 n={n} max-abs error vs FP64 oracle {err:.3e}")
        assert err < 1e-4


@pytest.mark.parametrize("k,tlog", [(13, 8), (12, 8), (11, 8), (10, 6), (9, 6)])
def test_complex64_tile_shapes(native, k, tlog):
    """Every complex64 forward tile with whole-tile transposes (32, 16 and 8
    amplitudes per thread on 256 threads, 16 and 8 on 64), forced by the
    tunables, against the FP64 oracle."""
    n = 20
    say(f"case complex64 tile shape n={n} k={k} tlog={tlog}")
    rng = np.random.default_rng(600 + k + tlog)
    gates = random_gates(rng, n, 250)
    c = build_circuit(n, gates, dtype="complex64")
    knobs = dict(tile_qubits=k, tile_threads_log2=tlog, half_lds=0)
    for key, v in knobs.items():
        N.tune_set(key, v)
    try:
        shape = c.tile_shape()
        c.forward()
    finally:
        for key in knobs:
            N.tune_set(key, -1)
    assert (shape["tile_qubits"], shape["threads_log2"]) == (k, tlog), shape
    err = np.max(np.abs(c.retrieve_amplitudes_on_host() - oracle.run_gates(n, gates).amp))
    say(f"  passes={c.plan_stats()[0]} max-abs error vs FP64 oracle {err:.3e}")
    assert err < 1e-4


@pytest.mark.parametrize("k,tlog", [(10, 8), (10, 7), (12, 9)])
def test_batch_tile_shapes(native, k, tlog):
    """The batched kernel's other tile shapes (tunables): 4 amplitudes per
    thread and state on 256 threads, 8 on 128 and 8 on 512, on circuits on
    both sides of the tile size (LDS path below it, tile path from it up),
    against the default shape in complex128 and, for complex64, against the
    complex128 results."""
    say(f"case batch shape k={k} tlog={tlog}")
    from damavand_amd import batch_forward_adjoint
    circuits = wl.batch_workload(16, 4500 + k + tlog, n_range=(8, 14), layer_range=(2, 3))
    e_ref, g_ref = batch_forward_adjoint(circuits, dtype="complex128")
    N.tune_set("batch_tile_qubits", k)
    N.tune_set("batch_threads_log2", tlog)
    try:
        e128, g128 = batch_forward_adjoint(circuits, dtype="complex128")
        e64, g64 = batch_forward_adjoint(circuits, dtype="complex64")
    finally:
        N.tune_set("batch_tile_qubits", -1)
        N.tune_set("batch_threads_log2", -1)
    worst = {"complex128": [0.0, 0.0], "complex64": [0.0, 0.0]}
    for i in range(len(circuits)):
        for key, e, g in (("complex128", e128, g128), ("complex64", e64, g64)):
            worst[key][0] = max(worst[key][0], abs(e[i] - e_ref[i]))
            if g_ref[i].size:
                worst[key][1] = max(worst[key][1], float(np.max(np.abs(g[i] - g_ref[i]))))
    say(f"  against the default shape: complex128 energy {worst['complex128'][0]:.3e} gradient "
        f"{worst['complex128'][1]:.3e}; complex64 energy {worst['complex64'][0]:.3e} gradient "
        f"{worst['complex64'][1]:.3e}")
    assert max(worst["complex128"]) <= 1e-12
    assert max(worst["complex64"]) <= 1e-5


def test_batch_streaming(native):
    """Streaming loads and stores of the batched kernel's scratch states
    (the default from 512 MiB of scratch; forced here on a small batch)
    change only the cache policy: bit-identical energies and gradients."""
    say("case batch streaming")
    from damavand_amd import batch_forward_adjoint
    circuits = wl.batch_workload(24, 4600, n_range=(8, 14), layer_range=(2, 3))
    for dtype in ("complex128", "complex64"):
        out = []
        for nt in (0, 1):
            N.tune_set("nontemporal", nt)
            try:
                out.append(batch_forward_adjoint(circuits, dtype=dtype))
            finally:
                N.tune_set("nontemporal", -1)
        (e0, g0), (e1, g1) = out
        same = np.array_equal(e0, e1) and all(np.array_equal(a, b) for a, b in zip(g0, g1))
        say(f"  {dtype}: bit-identical {same}")
        assert same


def test_batch_matches_single_circuits(native):
    say("case batch")
    circuits = wl.batch_workload(24, 4242, n_range=(3, 14), layer_range=(1, 4))
    from damavand_amd import batch_forward_adjoint
    for dtype, tol in (("complex128", 1e-10), ("complex64", 1e-3)):
        energy, grads = batch_forward_adjoint(circuits, dtype=dtype)
        worst_e = worst_g = 0.0
        for i, (n, gates, a, b) in enumerate(circuits):
            c = build_circuit(n, gates)
            e, g = c.gradient(a, wl.chain_pairs(n), b)
            worst_e = max(worst_e, abs(energy[i] - e))
            worst_g = max(worst_g, float(np.max(np.abs(grads[i] - g))) if g.size else 0.0)
        say(f"  {dtype}: energy {worst_e:.3e} gradient {worst_g:.3e} vs single-circuit adjoint")
        assert worst_e <= tol and worst_g <= tol
    # a few against the oracle's parameter shift
    for n, gates, a, b in circuits[:3]:
        energy, grads = batch_forward_adjoint([(n, gates, a, b)])
        g_ref = _param_shift(n, gates, a, wl.chain_pairs(n), b)
        assert np.max(np.abs(grads[0] - g_ref)) <= 1e-10


def test_hipgraph_replays(native):
    """Forward replayed from a captured hipGraph equals the eager forward on
    every replay, with eager work in between."""
    n = 21
    gates = wl.layered_circuit(n, 2, 8)
    ref = build_circuit(n, gates)
    ref.forward()
    want = ref.retrieve_amplitudes_on_host()
    c = build_circuit(n, gates)
    N.tune_set("graph", 1)
    try:
        for rep in range(4):
            c.reset_amplitudes()
            c.forward()
            got = c.retrieve_amplitudes_on_host()
            assert np.max(np.abs(got - want)) <= TOL, rep
            c.state.norm_sqr()  # eager work between replays
    finally:
        N.tune_set("graph", -1)


def _read_device(ptr, nbytes):
    """Raw bytes of device memory, copied with the HIP runtime the library
    itself is bound to."""
    memcpy = hip_memcpy()
    out = np.empty(nbytes, dtype=np.uint8)
    N.synchronize()
    assert memcpy(out.ctypes.data, ptr, nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return out


def test_states_of_other_layouts(native):
    """A circuit runs on states created under different layout settings (its
    plan is rebuilt for each), and operations that pair two states reject
    states whose element orders differ."""
    n = 25
    say(f"case states of other layouts n={n}")
    N.tune_set("page_fold_span_bits", 0)
    try:
        folded = StateVector(n)
    finally:
        N.tune_set("page_fold_span_bits", -1)
    plain = StateVector(n)
    gates = [(wl.GATE_H, -1, q, ()) for q in range(n)] + [(wl.GATE_CNOT, 3, n - 1, ()),
                                                         (wl.GATE_RY, -1, n - 2, (0.4,)), (wl.GATE_CZ, n - 1, 1, ())]
    c = build_circuit(n, gates)
    got = []
    for sv in (folded, plain, folded):
        sv.reset(0)
        N.call("dmv_circuit_forward", c._c.h, sv.handle, 0)
        got.append(sv.amplitudes())
    ref = oracle.run_gates(n, gates).amp
    for g in got:
        assert np.max(np.abs(g - ref)) <= TOL
    out = np.zeros(2)
    with pytest.raises(N.DamavandError, match="element order"):
        N.call("dmv_inner_product", folded.handle, plain.handle, N.dptr(out))
    with pytest.raises(N.DamavandError, match="element order"):
        N.call("dmv_state_copy", plain.handle, folded.handle)


@pytest.mark.parametrize("dtype,n,spread,fold", [("complex128", 19, False, False), ("complex64", 20, False, False),
                                                 ("complex128", 25, False, True), ("complex64", 26, False, True),
                                                 ("complex128", 25, True, True), ("complex64", 26, True, False)])
def test_state_layout_swizzled(native, dtype, n, spread, fold):
    """States past 2 MiB of amplitudes have the swizzled order: the raw device
    buffer holds amplitude i at the element the header's formula names, and
    loading, fused and per-gate passes over high qubits, reductions and
    readback all agree with the oracle.  With fold, the span-index fold (by
    default from 32 qubits on) is switched on for these sizes, where the span
    index is wide enough for it to act.  With spread, the size threshold of
    spread buffers is lowered so that these states are eight physical pieces
    mapped back to back."""
    N.tune_set("spread_min_bytes", (64 << 20) if spread else (1 << 62))
    N.tune_set("page_fold_span_bits", 0 if fold else 99)
    try:
        _check_state_layout(dtype, n, spread, fold)
    finally:
        N.tune_set("spread_min_bytes", -1)
        N.tune_set("page_fold_span_bits", -1)


def _check_state_layout(dtype, n, spread, fold):
    from tests.tile_emulator import element_map

    say(f"case state layout {dtype} n={n} spread={spread} fold={fold}")
    c64 = dtype == "complex64"
    tol = 1e-5 if c64 else TOL
    sv = StateVector(n, dtype)
    pieces = ctypes.c_int(0)
    N.call("dmv_state_pieces", sv.handle, ctypes.byref(pieces))
    assert pieces.value == (8 if spread else 1)
    fold_on = fold
    shift, blk, fold = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(-1)
    N.call("dmv_state_layout", sv.handle, ctypes.byref(shift), ctypes.byref(blk), ctypes.byref(fold))
    assert (shift.value, blk.value) == ((18, 7) if c64 else (17, 6))
    assert fold.value == (6 if fold_on else 0)
    ptr, nbytes = ctypes.c_void_p(), ctypes.c_uint64(0)
    N.call("dmv_state_device_ptr", sv.handle, ctypes.byref(ptr), ctypes.byref(nbytes))
    assert nbytes.value == (1 << n) * (8 if c64 else 16)
    rng = np.random.default_rng(n)
    amp = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    amp /= np.linalg.norm(amp)
    sv.load(amp)
    np.testing.assert_allclose(sv.amplitudes(), amp, atol=1e-7 if c64 else 0)
    raw = _read_device(ptr.value, nbytes.value).view(np.float32 if c64 else np.float64)
    raw = raw[0::2] + 1j * raw[1::2]
    want = amp.astype(np.complex64) if c64 else amp
    assert np.array_equal(raw[element_map(np.arange(1 << n), shift.value, blk.value, fold.value)], want)
    gates = [(wl.GATE_H, -1, n - 1, ()), (wl.GATE_CNOT, n - 1, n - 2, ()), (wl.GATE_RX, -1, n - 3, (0.7,)),
             (wl.GATE_CNOT, 0, n - 1, ()), (wl.GATE_H, -1, 1, ()), (wl.GATE_RZ, -1, n - 1, (1.1,))]
    ref = oracle.OracleState(n)
    ref.amp[:] = amp
    for kind, ctl, t, p in gates:
        ref.apply_gate(kind, t, ctl, p)
    for per_gate in (False, True):
        c = build_circuit(n, gates, dtype)
        N.call("dmv_state_pieces", c.state.handle, ctypes.byref(pieces))
        assert pieces.value == (8 if spread else 1)
        c.state.load(amp)
        c.forward(per_gate=per_gate)
        err = np.max(np.abs(c.retrieve_amplitudes_on_host() - ref.amp))
        say(f"  per_gate={per_gate} max-abs error {err:.3e}")
        assert err <= tol
        np.testing.assert_allclose(c.state.probabilities(), np.abs(ref.amp) ** 2, atol=tol)
        np.testing.assert_allclose(c.state.expect_z(), ref.expect_z(), atol=tol * 10)


def test_errors(native):
    c = Circuit(3, apply_method="gpu")
    with pytest.raises(N.DamavandError):
        c.add_hadamard_gate(3)
    with pytest.raises(N.DamavandError):
        c.add_cnot_gate(1, 1)
    st = StateVector(3)
    with pytest.raises(N.DamavandError):
        st.apply_matrix(np.eye(2), 5)
    with pytest.raises(N.DamavandError):
        st.apply_matrix(np.eye(2), 1, controls=[1])


def test_torch_autograd_single_and_batched(native):
    """The torch bridges: expectation() of one circuit and batch_expectation()
    of a packed batch, energies and theta.grad against the oracle (energy and
    parameter shift)."""
    import torch
    from damavand_amd.autograd import batch_expectation, expectation
    from damavand_amd.batch import pack_circuits
    say("case torch autograd")
    n = 8
    gates, a, b = wl.ansatz_circuit(n, 3, 77)
    pairs = wl.chain_pairs(n)
    c = build_circuit(n, gates)
    theta = torch.tensor([g[3][0] for g in gates if g[0] in wl.PARAMETRIC], dtype=torch.float64, requires_grad=True)
    e = expectation(c, theta, a, pairs, b)
    e.backward()
    st = oracle.run_gates(n, gates)
    assert abs(e.item() - st.expect_zz(a, pairs, b)) <= 1e-11
    pidx = [i for i, g in enumerate(gates) if g[0] in wl.PARAMETRIC]
    ref = []
    for gi in pidx:
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            g2 = list(gates)
            k, ctl, t, p = g2[gi]
            g2[gi] = (k, ctl, t, (p[0] + s,))
            vals.append(oracle.run_gates(n, g2).expect_zz(a, pairs, b))
        ref.append(0.5 * (vals[0] - vals[1]))
    np.testing.assert_allclose(theta.grad.numpy(), ref, atol=1e-10)

    circuits = wl.batch_workload(6, 88, n_range=(9, 13), layer_range=(2, 3))  # LDS and tile paths
    packed = pack_circuits(circuits)
    rng = np.random.default_rng(5)
    th = torch.tensor(rng.uniform(0, 2 * math.pi, packed.total_params), requires_grad=True)
    w = torch.tensor(rng.normal(size=packed.count))
    energies = batch_expectation(packed, th)
    (energies * w).sum().backward()
    # per circuit: energy and gradient at the substituted angles, on the oracle
    off = 0
    for i, (ni, g0, ai, bi) in enumerate(circuits):
        prm = th.detach().numpy()[off: off + packed.n_params[i]]
        it = iter(prm)
        gi = [(k, ctl, t, (next(it),) if k in wl.PARAMETRIC else p) for (k, ctl, t, p) in g0]
        chain = wl.chain_pairs(ni)
        ref_e = oracle.run_gates(ni, gi).expect_zz(ai, chain, bi)
        assert abs(energies[i].item() - ref_e) <= 1e-10
        # two angles of each circuit by parameter shift
        for j in (0, len(prm) - 1):
            pos = [q for q, g in enumerate(gi) if g[0] in wl.PARAMETRIC][j]
            vals = []
            for s in (0.5 * math.pi, -0.5 * math.pi):
                g2 = list(gi)
                k, ctl, t, p = g2[pos]
                g2[pos] = (k, ctl, t, (p[0] + s,))
                vals.append(oracle.run_gates(ni, g2).expect_zz(ai, chain, bi))
            assert abs(th.grad[off + j].item() - w[i].item() * 0.5 * (vals[0] - vals[1])) <= 1e-9
        off += packed.n_params[i]
