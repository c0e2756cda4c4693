"""Committed fixtures (tests/golden/golden.npz, written by make_golden.py):
the oracle and the workload generator on the CPU, the HIP path on the GPU."""
import os

import numpy as np
import pytest

import oracle
from damavand_amd import workloads as wl

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


def gates_of(name):
    k, c, t, p = (GOLDEN[name + s] for s in ("_kind", "_ctrl", "_tgt", "_prm"))
    return int(GOLDEN[name + "_n"]), [(int(a), int(b), int(d), tuple(e)) for a, b, d, e in zip(k, c, t, p)]


@pytest.mark.parametrize("name", ["circ_a", "circ_b"])
def test_oracle_reproduces_fixture(name):
    n, gates = gates_of(name)
    np.testing.assert_allclose(oracle.run_gates(n, gates).amp, GOLDEN[name + "_amp"], atol=1e-13)


def test_oracle_reference_known_answer():
    st = oracle.run_gates(2, [(wl.GATE_H, -1, 0, ()), (wl.GATE_H, -1, 1, ()), (wl.GATE_CNOT, 0, 1, ())])
    np.testing.assert_allclose(st.amp, GOLDEN["kat_2q"], atol=1e-15)


def test_generator_is_stable():
    g1 = wl.layered_circuit(20, 10, 1001)
    np.testing.assert_array_equal([g[3][0] for g in g1 if g[0] in wl.PARAMETRIC], GOLDEN["config1_angles"])
    g3, a, b = wl.ansatz_circuit(24, 20, 1003)
    np.testing.assert_array_equal([g[3][0] for g in g3 if g[0] in wl.PARAMETRIC], GOLDEN["config3_angles"])
    np.testing.assert_array_equal(a, GOLDEN["config3_a"])
    np.testing.assert_array_equal(b, GOLDEN["config3_b"])


def test_oracle_sampler_known_answers():
    np.testing.assert_array_equal(oracle.sample_uniforms(GOLDEN["sampler_p"], GOLDEN["sampler_u"]),
                                  GOLDEN["sampler_idx"])
    p2 = np.abs(GOLDEN["sampler2_amp"]) ** 2
    np.testing.assert_array_equal(oracle.sample_uniforms(p2, GOLDEN["sampler2_u"]), GOLDEN["sampler2_idx"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["circ_a", "circ_b"])
def test_gpu_reproduces_fixture(native, name):
    from damavand_amd.circuit import Circuit
    n, gates = gates_of(name)
    for per_gate in (False, True):
        c = Circuit(n, apply_method="gpu")
        for g in gates:
            c.add_gate(*g)
        c.forward(per_gate=per_gate)
        assert np.max(np.abs(c.retrieve_amplitudes_on_host() - GOLDEN[name + "_amp"])) <= 1e-12


@pytest.mark.gpu
def test_gpu_sampler_known_answers(native):
    from damavand_amd.circuit import StateVector
    st = StateVector(3)
    st.load(GOLDEN["sampler2_amp"])
    np.testing.assert_array_equal(st.sample_uniforms(GOLDEN["sampler2_u"]), GOLDEN["sampler2_idx"])
