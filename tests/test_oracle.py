"""Pins the C oracle (oracle/oracle.c) before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement (a
transposed matrix, a wrong sign, a wrong partner index, an off-by-one in the
sampler) fails at least one of them.  CPU only.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from oracle import brute
from damavand_amd import workloads as wl

X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)


@pytest.mark.parametrize("theta", [0.0, 0.3, 1.7, math.pi, 4.2, 2 * math.pi - 1e-3])
def test_rotations_equal_matrix_exponentials(theta):
    # gates.rs:418-439 / :490-511 / :562-577 are exp(-i theta P / 2)
    for kind, P in ((oracle.GATE_RX, X), (oracle.GATE_RY, Y), (oracle.GATE_RZ, Z)):
        got = oracle.gate_matrix(kind, [theta])
        np.testing.assert_allclose(got, expm(-0.5j * theta * P), atol=1e-15)


def test_fixed_gates_textbook():
    s = 1 / math.sqrt(2)
    np.testing.assert_allclose(oracle.gate_matrix(oracle.GATE_H), s * np.array([[1, 1], [1, -1]]), atol=0)
    np.testing.assert_array_equal(oracle.gate_matrix(oracle.GATE_X), X)
    np.testing.assert_array_equal(oracle.gate_matrix(oracle.GATE_Y), Y)
    np.testing.assert_array_equal(oracle.gate_matrix(oracle.GATE_Z), Z)
    np.testing.assert_array_equal(oracle.gate_matrix(oracle.GATE_S), np.diag([1, 1j]))
    np.testing.assert_allclose(oracle.gate_matrix(oracle.GATE_T), np.diag([1, np.exp(0.25j * math.pi)]), atol=1e-16)
    np.testing.assert_array_equal(oracle.gate_matrix(oracle.GATE_CNOT), X)


def test_u3_decomposition():
    rng = np.random.default_rng(3)
    for _ in range(20):
        th, ph, la = rng.uniform(0, 2 * math.pi, 3)
        u3 = oracle.gate_matrix(oracle.GATE_U3, [th, ph, la])
        rz = lambda a: expm(-0.5j * a * Z)
        ry = lambda a: expm(-0.5j * a * Y)
        ref = np.exp(0.5j * (ph + la)) * rz(ph) @ ry(th) @ rz(la)
        np.testing.assert_allclose(u3, ref, atol=1e-14)
        np.testing.assert_allclose(u3 @ u3.conj().T, np.eye(2), atol=1e-14)


def test_reference_known_answer_two_qubits():
    # circuit.rs:802-837: H(0), H(1), CNOT(0,1) on 2 qubits -> all amplitudes 1/2
    st = oracle.run_gates(2, [(oracle.GATE_H, -1, 0, ()), (oracle.GATE_H, -1, 1, ()),
                              (oracle.GATE_CNOT, 0, 1, ())])
    np.testing.assert_allclose(st.amp, np.full(4, 0.5), atol=1e-15)


def _random_gates(rng, n, count):
    gates = []
    kinds = [oracle.GATE_H, oracle.GATE_X, oracle.GATE_Y, oracle.GATE_Z, oracle.GATE_RX,
             oracle.GATE_RY, oracle.GATE_RZ, oracle.GATE_S, oracle.GATE_T, oracle.GATE_CNOT,
             oracle.GATE_U3, oracle.GATE_CZ]
    for _ in range(count):
        k = kinds[rng.integers(len(kinds))]
        t = int(rng.integers(n))
        c = -1
        if k in (oracle.GATE_CNOT, oracle.GATE_CZ):
            c = int(rng.integers(n - 1))
            if c >= t:
                c += 1
        p = tuple(rng.uniform(0, 2 * math.pi, 3))
        gates.append((k, c, t, p))
    return gates


def _brute_ops(gates):
    ops = []
    for k, c, t, p in gates:
        m = oracle.gate_matrix(oracle.GATE_Z if k == oracle.GATE_CZ else k, p)
        ops.append((m, t, () if c < 0 else (c,)))
    return ops


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (5, 3), (7, 4), (8, 5)])
def test_oracle_matches_kron_brute_force(n, seed):
    rng = np.random.default_rng(seed)
    if n == 1:
        gates = [g for g in _random_gates(rng, 2, 40) if g[1] < 0]
        gates = [(k, c, 0, p) for k, c, t, p in gates]
    else:
        gates = _random_gates(rng, n, 60)
    st = oracle.run_gates(n, gates)
    ref = brute.run(n, _brute_ops(gates))
    np.testing.assert_allclose(st.amp, ref, atol=1e-13, rtol=0)


def test_multi_control_matches_brute_force():
    rng = np.random.default_rng(11)
    n = 6
    st = oracle.OracleState(n)
    ops = []
    for _ in range(30):
        t = int(rng.integers(n))
        others = [q for q in range(n) if q != t]
        cs = tuple(int(q) for q in rng.choice(others, size=int(rng.integers(0, 3)), replace=False))
        m = oracle.gate_matrix(oracle.GATE_U3, rng.uniform(0, 6, 3))
        mask = 0
        for c in cs:
            mask |= 1 << c
        st.apply_matrix_mask(m, t, mask)
        ops.append((m, t, cs))
    np.testing.assert_allclose(st.amp, brute.run(n, ops), atol=1e-13)


def test_planted_transpose_is_caught():
    # a transposed gate matrix must be detected by the brute-force comparison
    rng = np.random.default_rng(2)
    n = 4
    gates = [g for g in _random_gates(rng, n, 30)]
    st = oracle.OracleState(n)
    for k, c, t, p in gates:
        m = oracle.gate_matrix(oracle.GATE_Z if k == oracle.GATE_CZ else k, p)
        st.apply_matrix(m.T, t, c)
    ref = brute.run(n, _brute_ops(gates))
    assert np.max(np.abs(st.amp - ref)) > 1e-3


def test_planted_wrong_target_is_caught():
    n = 3
    gates = [(oracle.GATE_RY, -1, 0, (0.7,)), (oracle.GATE_CNOT, 0, 2, ()), (oracle.GATE_RX, -1, 1, (1.1,))]
    st = oracle.OracleState(n)
    for k, c, t, p in gates:
        st.apply_gate(k, (t + 1) % n, c, p)
    assert np.max(np.abs(st.amp - brute.run(n, _brute_ops(gates)))) > 1e-3


@pytest.mark.parametrize("n,layers", [(6, 1), (12, 10), (14, 5)])
def test_rx_layer_workload_closed_form(n, layers):
    # examples/single_node_cpu/single_node.py:6-12 workload: product state
    st = oracle.run_gates(n, wl.rx_layers_circuit(n, layers))
    idx = np.arange(1 << n)
    np.testing.assert_allclose(st.amp, wl.rx_layers_closed_form(n, layers, idx), atol=1e-13)
    ez = st.expect_z()
    np.testing.assert_allclose(ez, [math.cos(layers * math.pi / (i + 1)) for i in range(n)], atol=1e-13)


def test_ghz_and_uniform_probabilities():
    n = 10
    gates = [(oracle.GATE_H, -1, 0, ())] + [(oracle.GATE_CNOT, i, i + 1, ()) for i in range(n - 1)]
    p = oracle.run_gates(n, gates).probabilities()
    exp = np.zeros(1 << n)
    exp[0] = exp[-1] = 0.5
    np.testing.assert_allclose(p, exp, atol=1e-15)
    p = oracle.run_gates(n, [(oracle.GATE_H, -1, q, ()) for q in range(n)]).probabilities()
    np.testing.assert_allclose(p, np.full(1 << n, 2.0 ** -n), atol=1e-16)
    assert abs(p.sum() - 1) < 1e-13


def test_cnot_permutes_basis_states():
    n = 5
    for x in [0, 3, 7, 21, 31]:
        for c, t in [(0, 1), (4, 0), (2, 3)]:
            st = oracle.OracleState(n)
            for q in range(n):
                if (x >> q) & 1:
                    st.apply_gate(oracle.GATE_X, q)
            st.apply_gate(oracle.GATE_CNOT, t, c)
            y = x ^ ((((x >> c) & 1)) << t)
            exp = np.zeros(1 << n, dtype=complex)
            exp[y] = 1
            np.testing.assert_array_equal(st.amp, exp)


def test_expect_zz_matches_kron_diagonal():
    rng = np.random.default_rng(5)
    n = 7
    gates = _random_gates(rng, n, 50)
    st = oracle.run_gates(n, gates)
    a = rng.normal(size=n)
    # a reversed pair and a repeated qubit (Z_3 Z_3 = I adds +b) as well
    pairs = [(i, i + 1) for i in range(n - 1)] + [(0, 5), (5, 0), (3, 3)]
    b = rng.normal(size=len(pairs))
    H = np.zeros((1 << n, 1 << n), dtype=complex)
    for i in range(n):
        H += a[i] * brute.full_operator(n, Z, i)
    for (i, j), bb in zip(pairs, b):
        H += bb * brute.full_operator(n, Z, i) @ brute.full_operator(n, Z, j)
    ref = np.real(np.vdot(st.amp, H @ st.amp))
    assert abs(st.expect_zz(a, pairs, b) - ref) < 1e-12


def test_ry_expectation_closed_form():
    n = 9
    th = np.linspace(0.1, 6.0, n)
    st = oracle.run_gates(n, [(oracle.GATE_RY, -1, q, (th[q],)) for q in range(n)])
    np.testing.assert_allclose(st.expect_z(), np.cos(th), atol=1e-14)


# ---- sampler: utils.rs:258-277 ------------------------------------------------

def test_sampler_known_answers_exact_binary_fractions():
    p = np.array([0.125, 0.25, 0.125, 0.5])  # cum = [0, .125, .375, .5, 1]
    u = np.array([1e-300, 0.125, 0.1250001, 0.375, 0.4, 0.5, 0.5000001, 1.0])
    np.testing.assert_array_equal(oracle.sample_uniforms(p, u), [0, 0, 1, 1, 2, 2, 3, 3])


def test_sampler_skips_zero_probability_outcomes():
    p = np.array([0.0, 0.5, 0.0, 0.0, 0.5, 0.0])
    u = np.array([1e-9, 0.25, 0.5, 0.5 + 1e-12, 0.9, 1.0])
    np.testing.assert_array_equal(oracle.sample_uniforms(p, u), [1, 1, 1, 4, 4, 4])


def test_sampler_scales_by_total_mass():
    # xsi = u * cum.last(), so an unnormalised vector samples the same outcomes
    p = np.array([1.0, 3.0])
    np.testing.assert_array_equal(oracle.sample_uniforms(p, np.array([0.2, 0.25, 0.26, 1.0])), [0, 0, 1, 1])


def test_sampler_rebuild_variant_agrees():
    rng = np.random.default_rng(0)
    p = rng.random(300)
    p /= p.sum()
    u = wl.sample_uniforms(7, 200)
    np.testing.assert_array_equal(oracle.sample_uniforms(p, u), oracle.sample_uniforms(p, u, rebuild=True))


def test_sampler_chi_square():
    from tests.stat_helpers import chi2_pvalue
    n = 8
    rng = np.random.default_rng(1)
    st = oracle.run_gates(n, _random_gates(rng, n, 40))
    p = st.probabilities()
    k = 200_000
    s = oracle.sample_uniforms(p, wl.sample_uniforms(123, k))
    assert chi2_pvalue(s, p) > 1e-4


# ---- generator -------------------------------------------------------------

def test_splitmix64_published_vector():
    # splitmix64 seeded with 0: first outputs as published with the algorithm
    r = wl.SplitMix64(0)
    assert r.next() == 0xE220A8397B1DCDAF
    assert r.next() == 0x6E789E6AA1B965F4
    assert r.next() == 0x06C45D188009454F


def test_sample_uniform_stream_matches_sequential():
    r = wl.SplitMix64(99)
    seq = np.array([r.uniform_open0() for _ in range(50)])
    np.testing.assert_array_equal(wl.sample_uniforms(99, 50), seq)


def test_workload_sizes():
    assert len(wl.layered_circuit(20, 10, 1001)) == 990
    assert len(wl.layered_circuit(30, 5, 1002)) == 745
    g, a, b = wl.ansatz_circuit(24, 20, 1003)
    assert wl.count_parameters(g) == 960 and len(g) == 1440 and a.size == 24 and b.size == 23
    assert len(wl.stress_circuit(33, 20, 1005)) == 980
