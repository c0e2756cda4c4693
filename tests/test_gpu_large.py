"""GPU checks at the full workload sizes, through size-independent
properties: norm preservation, U followed by U^dagger returning to |0>, and
the closed form of the reference's RX-layer workload (a product state)."""
import math
import sys

import numpy as np
import pytest

from damavand_amd import workloads as wl
from damavand_amd.circuit import Circuit

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


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


def build(n, gates):
    c = Circuit(n, apply_method="gpu")
    for g in gates:
        c.add_gate(*g)
    return c


@pytest.mark.parametrize("n", [26, 30])
def test_layered_round_trip(native, n):
    say(f"case round trip layered n={n}")
    gates = wl.layered_circuit(n, 5, 1002)
    c = build(n, gates)
    c.forward()
    norm = c.state.norm_sqr()
    ez = c.expectation_z()
    say(f"  after U: |norm-1| {abs(norm - 1):.3e}, <Z> range [{ez.min():.3f}, {ez.max():.3f}]")
    assert abs(norm - 1.0) < 1e-11
    assert np.all(np.abs(ez) <= 1 + 1e-12)
    inv = build(n, inverse_gates(gates))
    inv.state.copy_from(c.state)
    inv.forward()
    ez = inv.expectation_z()
    err = np.max(np.abs(ez - 1.0))
    say(f"  after U^dag U: max |<Z_i> - 1| {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("n", [24, 30])
def test_rx_layers_closed_form(native, n):
    say(f"case RX layers closed form n={n}")
    layers = 10
    c = build(n, wl.rx_layers_circuit(n, layers))
    c.forward()
    ez = c.expectation_z()
    ref = np.array([math.cos(layers * math.pi / (i + 1)) for i in range(n)])
    err = np.max(np.abs(ez - ref))
    say(f"  <Z_i> max-abs error {err:.3e}")
    assert err < 1e-12
    if n <= 24:  # amplitudes against the closed form (256 MiB to the host)
        amp = c.retrieve_amplitudes_on_host()
        err = np.max(np.abs(amp - wl.rx_layers_closed_form(n, layers, np.arange(1 << n))))
        say(f"  amplitudes max-abs error {err:.3e}")
        assert err < 1e-12
