"""The autotuner on the GPU: it leaves the caller's state and the tunables
alone, its choice applies to the circuit and is read back by a fresh process
from DMV_TUNE_CACHE, and results stay at parity with the oracle under the
chosen shapes (forward, reverse sweep, batch)."""
import json
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from damavand_amd import _native as N
from damavand_amd import workloads as wl
from damavand_amd.batch import batch_autotune, batch_forward_adjoint, pack_circuits
from damavand_amd.circuit import Circuit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_FRESH = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
from damavand_amd import workloads as wl
from damavand_amd.circuit import Circuit
c = Circuit(20)
for g in wl.layered_circuit(20, 1, 9):
    c.add_gate(*g)
print(json.dumps({"fwd": c.tile_shape(), "adj": c.tile_shape(adjoint=True)}))
"""


def say(*a):
    print(*a, flush=True)


@pytest.fixture
def tune_cache(tmp_path, monkeypatch):
    """A fresh DMV_TUNE_CACHE file for the test, with the process's tuning
    table emptied before and after so that no other test sees its choices."""
    path = tmp_path / "tune.jsonl"
    monkeypatch.setenv("DMV_TUNE_CACHE", str(path))
    N.call("dmv_tune_forget")
    yield path
    monkeypatch.delenv("DMV_TUNE_CACHE")
    N.call("dmv_tune_forget")


def build(n, gates, dtype="complex128"):
    c = Circuit(n, apply_method="gpu", dtype=dtype)
    for g in gates:
        c.add_gate(*g)
    return c


def test_autotune_forward_and_adjoint_parity_and_cache(native, tune_cache):
    cache = tune_cache
    before = {k: N.tune_get(k) for k in ("tile_qubits", "max_gen_per_pass", "nontemporal")}
    n = 20
    gates = wl.layered_circuit(n, 3, 1212)
    c = build(n, gates[:40])
    c.forward()
    amp0 = c.retrieve_amplitudes_on_host()
    say("case autotune forward n=20")
    k = c.autotune()
    shape = c.tile_shape()
    say(f"  chosen {shape}")
    assert k == shape["tile_qubits"] and 9 <= k <= 12
    # the caller's state and the tunables are left as they were
    assert np.array_equal(c.retrieve_amplitudes_on_host(), amp0)
    assert {k2: N.tune_get(k2) for k2 in before} == before
    # parity under the chosen shape
    c2 = build(n, gates)
    c2.autotune()
    c2.forward()
    ref = oracle.run_gates(n, gates)
    err = float(np.max(np.abs(c2.retrieve_amplitudes_on_host() - ref.amp)))
    say(f"  forward under the tuned shape: max-abs error {err:.3e}")
    assert err <= 1e-12
    # the reverse sweep
    na = 16
    ag, a, b = wl.ansatz_circuit(na, 2, 77)
    ca = build(na, ag)
    ka = ca.autotune(adjoint=True)
    say(f"  adjoint chosen {ca.tile_shape(adjoint=True)}")
    assert ka == ca.tile_shape(adjoint=True)["tile_qubits"]
    pairs = wl.chain_pairs(na)
    e, g = ca.gradient(a, pairs, b)
    st = oracle.run_gates(na, ag)
    assert abs(e - st.expect_zz(a, pairs, b)) <= 1e-11
    pidx = [i for i, x in enumerate(ag) if x[0] in wl.PARAMETRIC]
    for j in (0, len(pidx) // 2, len(pidx) - 1):
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            g2 = list(ag)
            kind, ctl, t, p = g2[pidx[j]]
            g2[pidx[j]] = (kind, ctl, t, (p[0] + s,))
            vals.append(oracle.run_gates(na, g2).expect_zz(a, pairs, b))
        assert abs(g[j] - 0.5 * (vals[0] - vals[1])) <= 1e-10
    # a fresh process reads the choices back from the cache file
    lines = [json.loads(x) for x in cache.read_text().splitlines()]
    assert any(x["key"].endswith("/n20/c128/forward") for x in lines)
    r = subprocess.run([sys.executable, "-c", _FRESH, ROOT], env=dict(os.environ), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    fresh = json.loads(r.stdout.strip().splitlines()[-1])
    say(f"  fresh process: {fresh['fwd']}")
    last = [x for x in lines if x["key"].endswith("/n20/c128/forward")][-1]
    assert fresh["fwd"]["tile_qubits"] == last["tile_qubits"]
    assert fresh["fwd"]["threads_log2"] == last["threads_log2"]
    assert fresh["fwd"]["max_gen_per_pass"] == last["max_gen_per_pass"]
    assert fresh["fwd"]["nontemporal"] == last["nontemporal"]


def test_batch_autotune_parity(native, tune_cache):
    say("case batch autotune")
    circuits = wl.batch_workload(96, 4343, n_range=(8, 14), layer_range=(2, 4))
    packed = pack_circuits(circuits)
    k = batch_autotune(packed)
    say(f"  chosen tile qubits {k}")
    assert 10 <= k <= 12
    e, gr = batch_forward_adjoint(circuits)
    for i in (0, 1, 2, 95):
        n, gates, a, b = circuits[i]
        c = build(n, gates)
        e1, g1 = c.gradient(a, wl.chain_pairs(n), b)
        assert abs(e[i] - e1) <= 1e-10 and np.max(np.abs(gr[i] - g1)) <= 1e-10
        ref = oracle.run_gates(n, gates).expect_zz(a, wl.chain_pairs(n), b)
        assert abs(e[i] - ref) <= 1e-10
