"""CPU-only checks of libdamavand-gpu: it loads, exports every declared
symbol, and its host logic (gate list, fused-pass scheduler, argument
checking) is right.  No kernel is launched here."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from damavand_amd import _native as N
from damavand_amd import workloads as wl
from damavand_amd.circuit import Circuit, expectation_signs
from tests.libhelpers import NativeCircuit, header_functions


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    declared = header_functions()
    assert len(declared) >= 55
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    # and the Python binding covers exactly them
    assert set(declared) == set(N.exported_symbols())


def test_reference_compat_symbols_present():
    # the extern "C" block of circuit.rs:27-51 and circuit_gpu.rs:3-24
    for name in ["get_number_of_available_gpus", "print_timers", "get_memory_for_gpu", "init_quantum_state",
                 "measure_on_gpu", "retrieve_amplitudes_on_host", "apply_one_qubit_gate_gpu_local",
                 "apply_one_qubit_gate_gpu_distributed", "split_amplitudes_between_gpus",
                 "exchange_amplitudes_between_gpus"]:
        assert name in header_functions()
        assert hasattr(N.lib(), name)


def test_version_string():
    assert "gfx950" in N.version()


def _oracle_run_plan(n, plan):
    kind, target, ctrl, mats, _ = plan
    st = oracle.OracleState(n)
    for k, t, c, m in zip(kind, target, ctrl, mats):
        st.apply_matrix_mask(m, int(t), int(c))
    return st.amp


def _random_gates(rng, n, count, kinds=None):
    kinds = kinds or [wl.GATE_H, wl.GATE_X, wl.GATE_Y, wl.GATE_Z, wl.GATE_RX, wl.GATE_RY, wl.GATE_RZ, wl.GATE_S,
                      wl.GATE_T, wl.GATE_CNOT, wl.GATE_U3, wl.GATE_CZ]
    out = []
    for _ in range(count):
        k = kinds[rng.integers(len(kinds))]
        t = int(rng.integers(n))
        c = -1
        if k in (wl.GATE_CNOT, wl.GATE_CZ):
            c = int(rng.integers(n - 1))
            c += c >= t
        out.append((k, c, t, tuple(rng.uniform(0, 2 * math.pi, 3))))
    return out


@pytest.mark.parametrize("n,tile,seed", [(13, 0, 1), (14, 9, 2), (16, 12, 3), (16, 10, 4), (18, 11, 5), (15, 13, 6)])
def test_fused_plan_reproduces_circuit(n, tile, seed):
    """The scheduler's op sequence (merged matrices, reordered across passes)
    applied by the oracle equals the oracle run of the original gate list."""
    rng = np.random.default_rng(seed)
    gates = _random_gates(rng, n, 300)
    nc = NativeCircuit(n, gates)
    dtype = N.C64 if tile == 13 else N.C128
    plan = nc.plan_ops(dtype=dtype, tile_qubits=tile)
    ref = oracle.run_gates(n, gates).amp
    np.testing.assert_allclose(_oracle_run_plan(n, plan), ref, atol=1e-12)
    # every pass touches at most `tile` qubits, including the coalescing low qubits
    kind, target, ctrl, _, pas = plan
    k = tile or 12
    low = 4 if dtype == N.C64 else 3
    for p in np.unique(pas):
        sel = pas == p
        qs = set(range(low))
        for t, c in zip(target[sel], ctrl[sel]):
            qs.add(int(t))
            qs.update(q for q in range(n) if (int(c) >> q) & 1)
        assert len(qs) <= k


def test_plan_of_layered_workload():
    n = 22
    gates = wl.layered_circuit(n, 3, 77)
    nc = NativeCircuit(n, gates)
    plan = nc.plan_ops()
    np.testing.assert_allclose(_oracle_run_plan(n, plan), oracle.run_gates(n, gates).amp, atol=1e-12)
    passes, phases, ops, hbm = nc.plan_stats()
    assert passes < len(gates) / 8  # fusion: far fewer HBM passes than gates
    assert hbm == pytest.approx(2 * 16 * 2 ** n * passes)


def test_plan_detects_planted_reordering_mistake():
    """Swapping two non-commuting ops of the plan must change the result, so
    the plan comparison above would catch a scheduler that reorders wrongly."""
    n = 13
    gates = [(wl.GATE_H, -1, 12, ()), (wl.GATE_CNOT, 12, 0, ()), (wl.GATE_RY, -1, 0, (0.7,))]
    plan = NativeCircuit(n, gates).plan_ops()
    kind, target, ctrl, mats, pas = plan
    order = list(range(len(kind)))
    order[1], order[2] = order[2], order[1]
    bad = (kind[order], target[order], ctrl[order], mats[order], pas[order])
    ref = oracle.run_gates(n, gates).amp
    assert np.max(np.abs(_oracle_run_plan(n, bad) - ref)) > 1e-3


def test_scheduler_merges_single_qubit_runs():
    n = 14
    gates = wl.layered_circuit(n, 1, 5)
    kind, target, ctrl, mats, pas = NativeCircuit(n, gates).plan_ops()
    # H, RX, RY, RZ on each qubit fold into one op per qubit, CNOTs stay
    assert len(kind) == n + (n - 1)


def test_add_gate_argument_checks():
    h = NativeCircuit(3, [])
    p = np.zeros(3)
    with pytest.raises(N.DamavandError):
        N.call("dmv_circuit_add_gate", h.h, wl.GATE_H, -1, 3, N.dptr(p), 0, None)   # target out of range
    with pytest.raises(N.DamavandError):
        N.call("dmv_circuit_add_gate", h.h, wl.GATE_CNOT, 1, 1, N.dptr(p), 0, None)  # control == target
    with pytest.raises(N.DamavandError):
        N.call("dmv_circuit_add_gate", h.h, wl.GATE_H, -1, 0, N.dptr(p), 1, None)    # H is not an observable
    with pytest.raises(N.DamavandError):
        N.call("dmv_circuit_add_gate", h.h, 99, -1, 0, N.dptr(p), 0, None)
    p[0] = float("nan")
    with pytest.raises(N.DamavandError):
        N.call("dmv_circuit_add_gate", h.h, wl.GATE_RX, -1, 0, N.dptr(p), 0, None)
    import ctypes
    bad = ctypes.c_void_p()
    with pytest.raises(N.DamavandError):
        N.call("dmv_circuit_create", 34, ctypes.byref(bad))


def test_set_parameters_semantics():
    # circuit.rs:308-322: zip(parametrised gates, values), truncated to the shorter
    import ctypes
    gates = [(wl.GATE_RX, -1, 0, (0.1,)), (wl.GATE_H, -1, 1, ()), (wl.GATE_RZ, -1, 1, (0.2,)),
             (wl.GATE_RY, -1, 0, (0.3,))]
    nc = NativeCircuit(2, gates)
    c = ctypes.c_int(0)
    N.call("dmv_circuit_num_parameters", nc.h, ctypes.byref(c))
    assert c.value == 3
    v = np.array([1.0, 2.0])
    N.call("dmv_circuit_set_parameters", nc.h, N.dptr(v), 2)
    out = np.zeros(3)
    N.call("dmv_circuit_get_parameters", nc.h, N.dptr(out), 3)
    np.testing.assert_array_equal(out, [1.0, 2.0, 0.3])


def test_apply_method_validation():
    with pytest.raises(ValueError, match="Apply method not recognized"):
        Circuit(3, apply_method="element_wise")
    for m in ("brute_force", "shuffle", "multithreading", "distributed_cpu", "distributed_gpu"):
        with pytest.raises(NotImplementedError):
            Circuit(3, apply_method=m)


def test_no_device_fails_loudly():
    if N.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(N.NoDeviceError):
        Circuit(3, apply_method="gpu")


def test_expectation_signs_follow_reference():
    # circuit.rs:494-513: -1 if the observable's target bit is set, else +1,
    # observables in insertion order, Pauli type ignored
    samples = np.array([0, 1, 2, 3, 5])
    out = expectation_signs(samples, [0, 2, 0])
    np.testing.assert_array_equal(out, [[1, 1, 1], [-1, 1, -1], [1, 1, 1], [-1, 1, -1], [-1, -1, -1]])


def test_pack_circuits_offsets():
    """The flat arrays of the batch ABI: offsets, parameter rows, coefficients."""
    from damavand_amd.batch import pack_circuits
    from damavand_amd import workloads as wl
    circuits = wl.batch_workload(5, 77)
    b = pack_circuits(circuits)
    assert b.count == 5
    g = 0
    c = 0
    p = 0
    for i, (ni, gates, a, bb) in enumerate(circuits):
        assert b.n[i] == ni and b.gate_offset[i] == g and b.coef_offset[i] == c and b.param_offset[i] == p
        for j, (kind, ctl, tgt, prm) in enumerate(gates):
            assert (b.kinds[g + j], b.ctrls[g + j], b.tgts[g + j]) == (kind, ctl, tgt)
            np.testing.assert_array_equal(b.params[3 * (g + j): 3 * (g + j) + len(prm)], prm)
        np.testing.assert_array_equal(b.coef[c: c + ni], a)
        np.testing.assert_array_equal(b.coef[c + ni: c + 2 * ni - 1], bb)
        g += len(gates)
        c += 2 * ni - 1
        p += wl.count_parameters(gates)
    assert b.gate_offset[5] == g and b.total_params == p


def test_packed_batch_parameter_substitution():
    """with_parameters replaces exactly the RX/RY/RZ angles, in gradient order."""
    from damavand_amd.batch import pack_circuits
    from damavand_amd import workloads as wl
    circuits = wl.batch_workload(4, 78)
    b = pack_circuits(circuits)
    theta0 = b.parameters()
    assert theta0.size == b.total_params
    expect = [p[0] for _, gates, _, _ in circuits for (k, c, t, p) in gates if k in wl.PARAMETRIC]
    np.testing.assert_array_equal(theta0, expect)
    theta = np.arange(b.total_params, dtype=float) * 0.01
    b2 = b.with_parameters(theta)
    np.testing.assert_array_equal(b2.parameters(), theta)
    rows = np.isin(b.kinds, list(wl.PARAMETRIC))
    np.testing.assert_array_equal(b2.params.reshape(-1, 3)[~rows], b.params.reshape(-1, 3)[~rows])
    np.testing.assert_array_equal(b.parameters(), theta0)  # the original is untouched


def test_set_parameters_rejects_without_change():
    """A non-finite value anywhere in the array is rejected before any gate
    changes, so the parameters (and the cached plans) stay as they were."""
    gates = [(wl.GATE_RX, -1, 0, (0.1,)), (wl.GATE_RZ, -1, 1, (0.2,)), (wl.GATE_RY, -1, 0, (0.3,))]
    nc = NativeCircuit(2, gates)
    v = np.array([1.0, float("nan"), 3.0])
    with pytest.raises(N.DamavandError, match="non-finite parameter 1"):
        N.call("dmv_circuit_set_parameters", nc.h, N.dptr(v), 3)
    out = np.zeros(3)
    N.call("dmv_circuit_get_parameters", nc.h, N.dptr(out), 3)
    np.testing.assert_array_equal(out, [0.1, 0.2, 0.3])
    # values past the parametrised gates are not looked at (circuit.rs:308-322 zips)
    v = np.array([1.0, 2.0, 3.0, float("inf")])
    N.call("dmv_circuit_set_parameters", nc.h, N.dptr(v), 4)
    N.call("dmv_circuit_get_parameters", nc.h, N.dptr(out), 3)
    np.testing.assert_array_equal(out, [1.0, 2.0, 3.0])


def _batch_call(b, **override):
    import ctypes
    arrs = dict(n=b.n, gate_offset=b.gate_offset, coef_offset=b.coef_offset, param_offset=b.param_offset)
    arrs.update({k: np.ascontiguousarray(v, dtype=np.int32) for k, v in override.items()})
    energy = np.zeros(b.count)
    grad = np.zeros(max(1, b.total_params))
    return N.lib().dmv_batch_forward_adjoint(
        ctypes.c_int(b.count), ctypes.c_int(0), N.iptr(arrs["n"]), N.iptr(arrs["gate_offset"]), N.iptr(b.kinds),
        N.iptr(b.ctrls), N.iptr(b.tgts), N.dptr(b.params), N.iptr(arrs["coef_offset"]), N.dptr(b.coef),
        N.iptr(arrs["param_offset"]), N.dptr(energy), N.dptr(grad))


def test_batch_rejects_bad_offsets_on_host():
    """Offsets that would index outside the packed arrays or the gradient are
    refused before the device is touched (so this runs without a GPU)."""
    from damavand_amd.batch import pack_circuits
    b = pack_circuits(wl.batch_workload(4, 79, n_range=(3, 5), layer_range=(1, 2)))
    po = b.param_offset.copy()
    po[2] += 1                      # padded: would write past the gradient of the last circuit
    assert _batch_call(b, param_offset=po) == 1
    assert "param_offset[2]" in N.last_error()
    po = b.param_offset.copy()
    po[1] = -5
    assert _batch_call(b, param_offset=po) == 1
    go = b.gate_offset.copy()
    go[2] = go[1] - 1               # decreasing gate range
    assert _batch_call(b, gate_offset=go) == 1
    assert "gate_offset decreases" in N.last_error()
    co = b.coef_offset.copy()
    co[3] = -1
    assert _batch_call(b, coef_offset=co) == 1
    assert "coef_offset[3]" in N.last_error()


def test_batch_rejects_too_many_parameters_per_pass_on_host():
    """The batched kernel keeps per-wave gradient sums for at most 128 IP ops
    of one reverse pass in LDS; a circuit whose reverse sweep puts more in
    one pass (200 rotations of one qubit: one hoisted IP group) is refused on
    the host, before the device is touched."""
    from damavand_amd.batch import pack_circuits
    n = 12
    gates = [(wl.GATE_H, -1, q, ()) for q in range(n)] + [(wl.GATE_RY, -1, 0, (0.01 * k,)) for k in range(200)]
    b = pack_circuits([(n, gates, np.ones(n), np.ones(n - 1))])
    assert _batch_call(b) == 1
    assert "parameters in one reverse pass" in N.last_error()


_TUNE_PROBE = r"""
import ctypes, json, sys
sys.path.insert(0, sys.argv[1])
from damavand_amd import _native as N
from damavand_amd import workloads as wl
from tests.libhelpers import NativeCircuit
out = {}
for n in (20, 21):
    c = NativeCircuit(n, wl.layered_circuit(n, 1, 5))
    for adj in (0, 1):
        v = [ctypes.c_int(0) for _ in range(4)]
        N.call("dmv_circuit_tile_shape", c.h, N.C128, adj, *[ctypes.byref(x) for x in v])
        out[f"{n}/{adj}"] = [x.value for x in v]
N.tune_set("tile_qubits", 10)
v = [ctypes.c_int(0) for _ in range(4)]
c = NativeCircuit(20, wl.layered_circuit(20, 1, 5))
N.call("dmv_circuit_tile_shape", c.h, N.C128, 0, *[ctypes.byref(x) for x in v])
out["20/0/tunable"] = [x.value for x in v]
print(json.dumps(out))
"""


def test_tune_cache_is_read_back_with_precedence(tmp_path):
    """The tuning table is filled from the DMV_TUNE_CACHE file on first use
    (later lines win, unrelated lines are skipped), applies to circuits of
    that size, and an explicit tunable takes precedence over it.  Runs in a
    fresh process (the table is read once per process); no device needed."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "tune.jsonl"
    cache.write_text(
        '{"key": "gfx950/n20/c128/forward", "tile_qubits": 12, "threads_log2": 9, "max_gen_per_pass": 6, '
        '"nontemporal": 0, "wgs_per_cu": 0, "ms": 3.0}\n'
        'not a cache line\n'
        '{"key": "gfx950/n20/c128/forward", "tile_qubits": 11, "threads_log2": 7, "max_gen_per_pass": 12, '
        '"nontemporal": 1, "wgs_per_cu": 0, "ms": 2.0}\n'
        '{"key": "gfx950/n20/c128/adjoint", "tile_qubits": 10, "threads_log2": 8, "max_gen_per_pass": 16, '
        '"nontemporal": 0, "wgs_per_cu": 0, "ms": 9.0}\n')
    env = dict(os.environ, DMV_TUNE_CACHE=str(cache))
    r = subprocess.run([sys.executable, "-c", _TUNE_PROBE, root], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    assert got["20/0"] == [11, 7, 12, 1]          # the later line for this key
    assert got["20/1"] == [10, 8, 16, 0]          # reverse sweeps never stream
    assert got["21/0"] == [12, 8, 8, 0]           # no entry: the defaults
    assert got["21/1"] == [11, 8, 8, 0]
    assert got["20/0/tunable"][0] == 10           # dmv_tune_set beats the table


def test_tunable_unset_restores_default():
    """A negative value unsets a tunable: dmv_tune_get reports -1 again and
    plans take the default (here the forward tile shape of a 20-qubit
    circuit)."""
    from tests.libhelpers import NativeCircuit
    c = NativeCircuit(20, wl.layered_circuit(20, 1, 5))

    def shape():
        v = [ctypes.c_int(0) for _ in range(4)]
        N.call("dmv_circuit_tile_shape", c.h, N.C128, 0, *[ctypes.byref(x) for x in v])
        return [x.value for x in v]

    default = shape()
    N.tune_set("tile_qubits", 10)
    try:
        assert N.tune_get("tile_qubits") == 10
        assert shape()[0] == 10
    finally:
        N.tune_set("tile_qubits", -1)
    assert N.tune_get("tile_qubits") == -1
    assert shape() == default
