"""The tile executor's plans, emulated on the CPU exactly as the kernels
index them (tests/tile_emulator.py), against the oracle."""
import math

import numpy as np
import pytest

import oracle
from damavand_amd import _native as N
from damavand_amd import workloads as wl
from tests.libhelpers import NativeCircuit
from tests.tile_emulator import TileEmulator, element_map, plan_json


def random_gates(rng, n, count):
    kinds = [wl.GATE_H, wl.GATE_X, wl.GATE_Y, wl.GATE_Z, wl.GATE_RX, wl.GATE_RY, wl.GATE_RZ, wl.GATE_S, wl.GATE_T,
             wl.GATE_CNOT, wl.GATE_CNOT, wl.GATE_CNOT, wl.GATE_U3, wl.GATE_CZ]
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


def zero_state(n):
    s = np.zeros(1 << n, dtype=complex)
    s[0] = 1
    return s


@pytest.mark.parametrize("n,k,tlog,gen,seed", [
    (13, 12, 8, 8, 1), (14, 12, 8, 4, 2), (15, 11, 8, 8, 3), (14, 9, 6, 8, 4), (15, 13, 9, 8, 5),
    (16, 12, 9, 40, 6), (14, 10, 8, 2, 7), (13, 12, 8, 100, 8), (14, 11, 6, 8, 9), (13, 10, 6, 16, 10)])
def test_forward_plan_emulation(n, k, tlog, gen, seed):
    rng = np.random.default_rng(seed)
    gates = random_gates(rng, n, 250)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=k, tlog=tlog, max_gen=gen)
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


def test_layered_workload_plan_emulation():
    n = 16
    gates = wl.layered_circuit(n, 3, 21)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


def test_complex64_plan_emulation():
    n = 15
    rng = np.random.default_rng(9)
    gates = random_gates(rng, n, 200)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, dtype=N.C64, tile_qubits=12, tlog=8, max_gen=8)
    got = TileEmulator(plan, elem=8).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


def test_plain_flag_marks_select_free_ops():
    """Most GEN ops of a layered circuit need no per-pair selects, and the flag says so exactly."""
    n = 16
    nc = NativeCircuit(n, wl.layered_circuit(n, 3, 5))
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    gens = [op for P in plan["passes"] for ph in P["phases"] for op in ph["ops"] if op["kind"] == 0]
    assert sum(op["plain"] for op in gens) >= 0.9 * len(gens)
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, nc.gates).amp, atol=1e-12)


def test_renaming_removes_register_cnots():
    """CNOT ladders among register qubits emit no device op (only the map changes)."""
    n = 14
    gates = [(wl.GATE_H, -1, q, ()) for q in range(n)] + [(wl.GATE_CNOT, q, q + 1, ()) for q in range(n - 1)]
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=100)
    swaps = sum(1 for P in plan["passes"] for ph in P["phases"] for op in ph["ops"] if op["kind"] == 2)
    assert swaps < n - 1
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


def test_planted_mask_mistake_is_caught():
    n = 13
    rng = np.random.default_rng(11)
    gates = random_gates(rng, n, 120)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=100)
    for P in plan["passes"]:
        for ph in P["phases"]:
            for op in ph["ops"]:
                if op["kind"] == 0:
                    op["fl"] ^= 1  # flip the member assignment of one pair
                    op["plain"] = 0
                    break
    got = TileEmulator(plan).forward(zero_state(n))
    assert np.max(np.abs(got - oracle.run_gates(n, gates).amp)) > 1e-6


@pytest.mark.parametrize("n,layers,k,tlog", [(13, 2, 11, 8), (14, 2, 12, 8), (13, 3, 9, 6)])
def test_adjoint_plan_emulation(n, layers, k, tlog):
    """The reverse sweep (G^dagger on psi and lambda, IP partials) emulated,
    against parameter shift on the oracle."""
    gates, a, b = wl.ansatz_circuit(n, layers, 500 + n)
    pairs = wl.chain_pairs(n)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, adjoint=True, tile_qubits=k, tlog=tlog, max_gen=8)
    st = oracle.run_gates(n, gates)
    idx = np.arange(1 << n)
    h = np.zeros(1 << n)
    for q in range(n):
        h += np.where((idx >> q) & 1, -a[q], a[q])
    for (i, j), bb in zip(pairs, b):
        h += np.where(((idx >> i) ^ (idx >> j)) & 1, -bb, bb)
    psi0, lam, grads = TileEmulator(plan).reverse(st.amp, h * st.amp)
    # psi is uncomputed back to |0>
    np.testing.assert_allclose(psi0, zero_state(n), atol=1e-12)
    pidx = [i for i, g in enumerate(gates) if g[0] in wl.PARAMETRIC]
    ref = []
    for gi in pidx:
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            g2 = list(gates)
            kk, c, t, p = g2[gi]
            g2[gi] = (kk, c, t, (p[0] + s,))
            vals.append(oracle.run_gates(n, g2).expect_zz(a, pairs, b))
        ref.append(0.5 * (vals[0] - vals[1]))
    np.testing.assert_allclose(grads, ref, atol=1e-10)


def test_gapped_layout_emulation():
    """A state past 2 MiB of amplitudes (swizzled order): the planner's element
    offsets and the kernels' element map of tile base | thread offset address
    exactly the state's elements (stores cover each once, results match the
    oracle)."""
    n = 19
    rng = np.random.default_rng(31)
    gates = random_gates(rng, n, 120) + [(wl.GATE_H, -1, q, ()) for q in range(12, n)]
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    assert plan["layout"] == [17, 6, 0]
    assert max(max(P["greg_ld"]) for P in plan["passes"]) >= (1 << 17)
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


def test_page_fold_emulation():
    """A state of more than 64 spans, where the span index folds: passes over
    the highest qubits (register offsets that cross the fold) address every
    element once and match the oracle."""
    n = 24
    gates = [(wl.GATE_H, -1, q, ()) for q in range(15, n)] + [(wl.GATE_CNOT, n - 1, 3, ()),
                                                               (wl.GATE_RY, -1, n - 2, (0.3,)),
                                                               (wl.GATE_CNOT, 2, n - 1, ())]
    nc = NativeCircuit(n, gates)
    N.tune_set("page_fold_span_bits", 0)  # fold from the smallest states on (default: 32 qubits and up)
    try:
        plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    finally:
        N.tune_set("page_fold_span_bits", -1)
    assert plan["layout"] == [17, 6, 6]
    # some register offset has a span part of 64 or more, so the fold acts on it
    assert any((g >> 17) >= 64 for P in plan["passes"] for g in P["greg_ld"])
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


@pytest.mark.parametrize("runs", [4, 5, 6])
def test_wider_coalescing_runs_emulation(runs):
    """Plans that keep more low qubits in lanes at load/store time."""
    n = 15
    rng = np.random.default_rng(20 + runs)
    gates = random_gates(rng, n, 200)
    nc = NativeCircuit(n, gates)
    N.tune_set("coalesce_qubits", runs)
    try:
        plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    finally:
        N.tune_set("coalesce_qubits", -1)
    for P in plan["passes"]:
        assert P["gthr_ld"][:runs] == list(range(runs)) and P["gthr_st"][:runs] == list(range(runs))
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


@pytest.mark.parametrize("shift,blk,fold", [(17, 6, 6), (18, 7, 6), (17, 6, 2), (17, 6, 0)])
def test_element_map_properties(shift, blk, fold):
    """The documented element map is a bijection, XOR-linear, and keeps 1 KiB
    blocks contiguous; with the fold, a tile over qubits 23-31 (33 qubits)
    touches spans of every value of the folded low span bits."""
    n = 24
    i = np.arange(1 << n, dtype=np.int64)
    e = element_map(i, shift, blk, fold)
    assert np.array_equal(np.sort(e), i)
    assert np.array_equal(e ^ (i & ((1 << blk) - 1)), element_map(i & ~np.int64((1 << blk) - 1), shift, blk, fold))
    rng = np.random.default_rng(shift + fold)
    a = rng.integers(0, 1 << 33, size=4096, dtype=np.int64)
    b = rng.integers(0, 1 << 33, size=4096, dtype=np.int64)
    assert np.array_equal(element_map(a ^ b, shift, blk, fold),
                          element_map(a, shift, blk, fold) ^ element_map(b, shift, blk, fold))
    hi = np.zeros(512, dtype=np.int64)
    for j in range(9):
        hi |= ((np.arange(512) >> j) & 1) << (23 + j + (shift - 17))
    low_span = (element_map(hi, shift, blk, fold) >> shift) & 63
    assert len(np.unique(low_span)) == (1 << fold if fold else 1)


@pytest.mark.parametrize("dtype,k,elem", [(N.C128, 12, 16), (N.C64, 13, 8)])
def test_split_transposes_emulation(dtype, k, elem):
    """Half-LDS passes: the planner keeps one local bit at the top thread bit
    of every layout of a pass, and each transpose runs as two rounds (the
    threads of either half) through half a tile of LDS.  Emulated round by
    round with the LDS cleared in between, against the oracle; with the
    tunable off, the same circuit plans full transposes."""
    n = 16
    rng = np.random.default_rng(41)
    gates = random_gates(rng, n, 300)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, dtype=dtype, tile_qubits=k, tlog=8, max_gen=8)
    split = [P for P in plan["passes"] if P["split"]]
    assert len(split) >= len(plan["passes"]) // 2
    assert any(P["transpose_st"] or any(ph["transpose_in"] for ph in P["phases"]) for P in split)
    for P in split:
        w = P["split"] - (4 if elem == 16 else 3)
        lays = [P["lay_ld"], P["lay_st"]] + [ph["lay"] for ph in P["phases"]]
        assert all(L["thr"][7] == w for L in lays)
        assert P["lds_half"] == 1
    got = TileEmulator(plan, elem=elem).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12 if elem == 16 else 1e-12)
    N.tune_set("half_lds", 0)
    try:
        full = plan_json(nc.h, dtype=dtype, tile_qubits=k, tlog=8, max_gen=8)
    finally:
        N.tune_set("half_lds", -1)
    assert all(P["split"] == 0 and P["lds_half"] == 0 for P in full["passes"])
    got = TileEmulator(full, elem=elem).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)


def test_planted_split_mistake_is_caught():
    """A split on a bit that is not the layouts' top thread bit moves amplitudes
    between the halves: the emulated rounds then read what their round did
    not write (or write outside half the LDS)."""
    n = 15
    rng = np.random.default_rng(43)
    gates = random_gates(rng, n, 200)
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    P = next(P for P in plan["passes"] if P["split"] and any(ph["transpose_in"] for ph in P["phases"]))
    P["split"] -= 1
    with pytest.raises(AssertionError):
        TileEmulator(plan).forward(zero_state(n))


def test_sign_flip_diagonals():
    """Z and CZ (exactly diag(1, -1)) are flagged plain and applied as sign
    flips; S and T keep the general diagonal; register and thread-bit
    targets and controls, against the oracle."""
    n = 14
    rng = np.random.default_rng(47)
    gates = [(wl.GATE_H, -1, q, ()) for q in range(n)]
    for _ in range(60):
        t = int(rng.integers(n))
        c = int(rng.integers(n - 1))
        c += c >= t
        k = [wl.GATE_CZ, wl.GATE_Z, wl.GATE_S, wl.GATE_T, wl.GATE_RY][int(rng.integers(5))]
        gates.append((k, c if k == wl.GATE_CZ else -1, t, (float(rng.uniform(0, 6.3)),)))
    nc = NativeCircuit(n, gates)
    plan = plan_json(nc.h, tile_qubits=12, tlog=8, max_gen=8)
    diags = [op for P in plan["passes"] for ph in P["phases"] for op in ph["ops"] if op["kind"] == 1]
    flips = [op for op in diags if op["plain"]]
    assert flips and len(flips) < len(diags)
    assert all(op["m"] == [1.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0] for op in flips)
    assert any(op["tbit"] < 64 for op in flips) and any(op["ctrl_thr"] for op in flips)
    got = TileEmulator(plan).forward(zero_state(n))
    np.testing.assert_allclose(got, oracle.run_gates(n, gates).amp, atol=1e-12)
