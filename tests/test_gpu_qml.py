"""The PennyLane-style facade (damavand_amd.qml) on the GPU against the
oracle: exact expval / var / probs / state of a 4-wire circuit with Pauli-X,
Y and Z observables, tensor products and a Hamiltonian; torch gradients of a
loss over all of them against the oracle's parameter-shift rule; and the
shot-based outputs against the exact values."""
import math

import numpy as np
import pytest
import torch

import oracle
from damavand_amd import qml
from damavand_amd import workloads as wl

pytestmark = pytest.mark.gpu

N_WIRES = 4
X0 = np.array([0.31, 1.7, -0.4, 2.2, 0.9, 1.1, -2.3, 0.6, 0.8])


def qml_circuit(x):
    qml.Hadamard(wires=0)
    qml.RX(x[0], wires=1)
    qml.RY(x[1], wires=2)
    qml.CNOT(wires=[0, 1])
    qml.Rot(x[2], x[3], x[4], wires=3)
    qml.CZ(wires=[1, 3])
    qml.U3(x[5], x[6], x[7], wires=2)
    qml.RZ(x[8], wires=0)
    qml.CNOT(wires=[2, 0])
    return (qml.expval(qml.PauliZ(0)), qml.expval(qml.PauliX(1)), qml.expval(qml.PauliY(2) @ qml.PauliZ(3)),
            qml.expval(qml.Hamiltonian([0.7, -1.3, 0.4], [qml.PauliZ(1) @ qml.PauliZ(2), qml.PauliX(0),
                                                          qml.PauliY(3)])),
            qml.var(qml.PauliZ(2)), qml.probs(wires=[3, 1]))


def oracle_gates(x):
    return [(wl.GATE_H, -1, 0, ()), (wl.GATE_RX, -1, 1, (x[0],)), (wl.GATE_RY, -1, 2, (x[1],)),
            (wl.GATE_CNOT, 0, 1, ()), (wl.GATE_RZ, -1, 3, (x[2],)), (wl.GATE_RY, -1, 3, (x[3],)),
            (wl.GATE_RZ, -1, 3, (x[4],)), (wl.GATE_CZ, 1, 3, ()), (wl.GATE_U3, -1, 2, (x[5], x[6], x[7])),
            (wl.GATE_RZ, -1, 0, (x[8],)), (wl.GATE_CNOT, 2, 0, ())]


PAULI = {"X": np.array([[0, 1], [1, 0]], dtype=complex), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0]).astype(complex)}


def pauli_expect(amp, ops):
    """<psi| prod_w P_w |psi> with qubit q = bit q of the index."""
    t = amp.reshape((2,) * N_WIRES)  # axis a is qubit N-1-a
    phi = t
    for w, p in ops.items():
        phi = np.moveaxis(np.tensordot(PAULI[p], phi, axes=([1], [N_WIRES - 1 - w])), 0, N_WIRES - 1 - w)
    return float(np.real(np.vdot(t, phi)))


def reference_outputs(x):
    amp = oracle.run_gates(N_WIRES, oracle_gates(x)).amp
    z2 = pauli_expect(amp, {2: "Z"})
    p = np.abs(amp) ** 2
    pm = np.zeros(4)
    for i in range(16):
        pm[((i >> 3) & 1) * 2 + ((i >> 1) & 1)] += p[i]  # wires [3, 1], wire 3 most significant
    return [pauli_expect(amp, {0: "Z"}), pauli_expect(amp, {1: "X"}), pauli_expect(amp, {2: "Y", 3: "Z"}),
            0.7 * pauli_expect(amp, {1: "Z", 2: "Z"}) - 1.3 * pauli_expect(amp, {0: "X"})
            + 0.4 * pauli_expect(amp, {3: "Y"}), 1.0 - z2 * z2, pm], amp


def test_exact_outputs_and_gradients(native):
    dev = qml.device("damavand.qubit", wires=N_WIRES, apply_method="gpu")
    node = qml.qnode(dev)(qml_circuit)
    x = torch.tensor(X0, requires_grad=True)
    out = node(x)
    ref, amp = reference_outputs(X0)
    for i, (o, r) in enumerate(zip(out, ref)):
        err = float(np.max(np.abs(o.detach().numpy() - r)))
        print(f"  output {i}: max-abs error {err:.2e}", flush=True)
        assert err <= 1e-12
    # the state, with U3's global phase
    st = qml.qnode(dev)(lambda v: (qml_circuit(v), qml.state())[1])(X0)
    assert np.max(np.abs(st - amp)) <= 1e-12
    # torch gradient of a loss over every output against the oracle's shift rule
    w_p = np.array([0.3, -1.0, 2.0, 0.5])
    loss = (out[0] + 2 * out[1] - out[2] + 0.5 * out[3] + 3 * out[4] + (out[5] * torch.tensor(w_p)).sum())
    loss.backward()
    want = np.zeros(X0.size)
    e_z2 = pauli_expect(amp, {2: "Z"})
    for i in range(X0.size):
        lin = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            xs = X0.copy()
            xs[i] += s
            r, a2 = reference_outputs(xs)
            lin.append(np.array([r[0] + 2 * r[1] - r[2] + 0.5 * r[3] + float(np.dot(r[5], w_p)),
                                 pauli_expect(a2, {2: "Z"})]))
        d = 0.5 * (lin[0] - lin[1])
        want[i] = d[0] + 3 * (-2.0 * e_z2 * d[1])  # var = 1 - <Z2>^2
    err = float(np.max(np.abs(x.grad.numpy() - want)))
    print(f"  gradient vs oracle parameter shift: {err:.2e}", flush=True)
    assert err <= 1e-10


def test_shot_outputs(native):
    shots = 20000
    dev = qml.device("damavand.qubit", wires=N_WIRES, shots=shots, seed=11)

    @qml.qnode(dev)
    def node(x):
        qml_circuit(x)
        return (qml.expval(qml.PauliZ(0)), qml.expval(qml.PauliX(1)), qml.probs(wires=[3, 1]),
                qml.sample(wires=[0, 2]), qml.sample(qml.PauliZ(3)))

    e0, e1, p, s, sz = node(X0)
    ref, _ = reference_outputs(X0)
    for got, r in ((e0, ref[0]), (e1, ref[1])):
        sd = math.sqrt(max(1.0 - r * r, 1e-12) / shots)
        assert abs(float(got) - r) < 5.5 * sd
    assert abs(p.sum() - 1.0) < 1e-12
    assert np.all(np.abs(p - ref[5]) < 5.5 * np.sqrt(ref[5] * (1 - ref[5]) / shots) + 1e-12)
    assert s.shape == (shots, 2) and set(np.unique(s)) <= {0, 1}
    assert sz.shape == (shots,) and set(np.unique(sz)) <= {-1.0, 1.0}
