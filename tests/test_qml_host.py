"""Host-side pieces of the PennyLane-style facade (damavand_amd.qml): tape
recording, observable algebra, marginal ordering and diagonal tables.  No
device needed."""
import numpy as np
import pytest

from damavand_amd import qml


def test_device_arguments():
    dev = qml.device("damavand.qubit", wires=["a", "b", 7], apply_method="gpu")
    assert dev.num_wires == 3 and dev.wire("b") == 1 and dev.wire(7) == 2
    with pytest.raises(ValueError):
        dev.wire("zz")
    with pytest.raises(NotImplementedError):
        qml.device("damavand.qubit", wires=2, apply_method="multithreading")
    with pytest.raises(ValueError):
        qml.device("default.qubit", wires=2)


def test_tape_records_gates_and_takes_observables_off_it():
    tape = qml._Tape()
    qml._ACTIVE.append(tape)
    try:
        qml.Hadamard(wires=0)
        qml.RX(0.3, wires=1)
        qml.CNOT(wires=[0, 1])
        m = qml.expval(qml.PauliZ(0) @ qml.PauliX(1))
        h = qml.expval(qml.Hamiltonian([0.5, -2.0], [qml.PauliZ(0), qml.PauliY(1)]))
    finally:
        qml._ACTIVE.pop()
    assert [o.name for o in tape.ops] == ["Hadamard", "RX", "CNOT"]
    assert m.obs.terms == [(1.0, {0: "Z", 1: "X"})]
    assert h.obs.terms == [(0.5, {0: "Z"}), (-2.0, {1: "Y"})]


def test_marginal_matches_brute_force():
    rng = np.random.default_rng(3)
    n = 5
    p = rng.random(1 << n)
    p /= p.sum()
    for wires in ([0], [3], [1, 4], [4, 1], [2, 0, 3]):
        got = qml._marginal(p, list(wires), n)
        want = np.zeros(1 << len(wires))
        for x in range(1 << n):
            j = 0
            for w in wires:  # wires[0] is the most significant bit
                j = (j << 1) | ((x >> w) & 1)
            want[j] += p[x]
        np.testing.assert_allclose(got, want, atol=1e-15)


def test_parity_tables_and_groups():
    obs = qml.Observable([(2.0, {0: "Z", 2: "Z"}), (-1.0, {1: "X"}), (0.5, {})])
    groups = {tuple(sorted(b.items())): (w, t) for b, w, t in qml._groups(obs)}
    w, t = groups[()]
    assert w == [0, 2]
    # j = bit0 + 2 bit2: Z0 Z2 = +1, -1, -1, +1, plus the identity's 0.5
    np.testing.assert_allclose(t, [2.5, -1.5, -1.5, 2.5])
    w, t = groups[((1, "X"),)]
    assert w == [1]
    np.testing.assert_allclose(t, [-1.0, 1.0])
