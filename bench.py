"""Benchmark of the statevector gate-application path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one fused forward of the 745-gate layered circuit (5 layers of H,
RX, RY, RZ on every qubit plus a CNOT ladder) over a 30-qubit complex128
state (16 GiB) that stays resident in HBM.  Prints one JSON line:

* value: HBM GB/s of the whole forward = algorithmic bytes of its fused
  passes (2 x 16 B x 2^30 per pass) / wall time per step;
* roofline: the fused tile kernel's achieved GB/s from its own HIP-event
  launch durations (same stream as the launches) against the 8 TB/s peak;
* cpu_baseline: the oracle (C/OpenMP restatement of the reference's
  multithreading path) timed on this host on a bounded sample of the same
  workload;
* adjoint: the 24-qubit, 960-parameter ansatz gradient by the adjoint
  kernels versus parameter shift with plain fused forwards, on the same GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "gate-apply HBM GB/s @30q c128; adjoint vs param-shift grads/sec (same GPU)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    from damavand_amd import _native as N
    N.synchronize()


def build_circuit(n, gates, dtype="complex128"):
    from damavand_amd.circuit import Circuit
    c = Circuit(n, apply_method="gpu", dtype=dtype)
    for g in gates:
        c.add_gate(*g)
    return c


def bench_forward(n, layers, steps, warmup):
    from damavand_amd import _native as N
    from damavand_amd import workloads as wl
    gates = wl.layered_circuit(n, layers, 1002)
    c = build_circuit(n, gates)
    passes, phases, ops, hbm_bytes = c.plan_stats()
    log(f"[bench] {n}q layered x{layers}: {len(gates)} gates -> {passes} passes, {phases} phases, {ops} ops")
    for _ in range(warmup):
        c.forward()
    sync()
    N.profile_reset()
    N.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.forward()
    sync()
    t1 = time.perf_counter()
    N.profile_enable(False)
    launches, ms_total, bytes_total = N.profile_get(N.KC_TILE)
    ms_step = (t1 - t0) * 1e3 / steps
    # fused kernel only, from its own launch events
    avg_ms = ms_total / max(1, launches)
    per_launch_bytes = bytes_total / max(1, launches)
    achieved = per_launch_bytes / (avg_ms * 1e6) if avg_ms > 0 else 0.0
    # sanity of the result
    norm = c.state.norm_sqr()
    pieces = ctypes.c_int(0)
    N.call("dmv_state_pieces", c.state.handle, ctypes.byref(pieces))
    log(f"[bench] forward {ms_step:.2f} ms/step, tile kernel {avg_ms:.3f} ms/launch, {achieved:.0f} GB/s, "
        f"|norm-1| {abs(norm - 1):.2e}")
    return {
        "gates": len(gates), "passes": passes, "phases": phases, "ops": ops, "hbm_bytes": hbm_bytes,
        "ms_step": ms_step, "launches": launches, "avg_launch_ms": avg_ms, "achieved": achieved,
        "per_launch_bytes": per_launch_bytes, "norm_err": abs(norm - 1), "state_pieces": pieces.value,
    }, c, gates


def bench_per_gate(c):
    from damavand_amd import _native as N
    c.reset_amplitudes()
    sync()
    N.profile_reset()
    N.profile_enable(True)
    t0 = time.perf_counter()
    c.forward(per_gate=True)
    sync()
    t1 = time.perf_counter()
    N.profile_enable(False)
    launches, ms_total, bytes_total = N.profile_get(N.KC_GATE)
    return {"gates": launches, "ms": (t1 - t0) * 1e3, "kernel_ms": ms_total,
            "GBps": bytes_total / (ms_total * 1e6) if ms_total > 0 else 0.0}


def bench_copy(nbytes):
    from damavand_amd import _native as N
    ms = N.copy_benchmark(nbytes, 5)
    return 2.0 * nbytes / (ms * 1e6)


def bench_adjoint(n, layers, shift_params):
    from damavand_amd import workloads as wl
    gates, a, b = wl.ansatz_circuit(n, layers, 1003)
    pairs = wl.chain_pairs(n)
    c = build_circuit(n, gates)
    e, g = c.gradient(a, pairs, b)  # warm-up (plans, buffers)
    sync()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        e, g = c.gradient(a, pairs, b)
    sync()
    t_adj = (time.perf_counter() - t0) / reps
    theta = c.get_parameters()
    # parameter shift with plain fused forwards on a seeded subset of parameters
    rng = np.random.default_rng(1003)
    idx = np.sort(rng.choice(theta.size, size=min(shift_params, theta.size), replace=False))
    ps = np.zeros(idx.size)
    c.reset_amplitudes()
    c.forward()
    sync()
    t0 = time.perf_counter()
    for j, k in enumerate(idx):
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            th = theta.copy()
            th[k] += s
            c.set_parameters(th)
            c.reset_amplitudes()
            c.forward()
            vals.append(c.expectation(a, pairs, b))
        ps[j] = 0.5 * (vals[0] - vals[1])
    sync()
    t_ps = time.perf_counter() - t0
    c.set_parameters(theta)
    per_param = t_ps / idx.size
    diff = float(np.max(np.abs(g[idx] - ps)))
    log(f"[bench] adjoint {n}q {theta.size} params: {t_adj * 1e3:.1f} ms; parameter shift "
        f"{per_param * 1e3:.2f} ms/param; max |adjoint - shift| {diff:.2e}")
    return {
        "workload": f"{n}-qubit RY/RZ + CNOT-ring ansatz, {layers} layers, {theta.size} parameters, Z/ZZ chain H",
        "adjoint_s": t_adj, "adjoint_grads_per_s": theta.size / t_adj,
        "param_shift_s_per_param": per_param, "param_shift_grads_per_s": 1.0 / per_param,
        "param_shift_sampled_params": int(idx.size),
        "param_shift_full_s_extrapolated": per_param * theta.size,
        "speedup": (per_param * theta.size) / t_adj, "max_abs_diff_vs_shift": diff, "energy": e,
    }


def bench_batch(count):
    """4096 ansatz circuits (12-18 qubits, 4-12 layers), forward + adjoint, one launch each dtype."""
    from damavand_amd import _native as N
    from damavand_amd.batch import batch_forward_adjoint_packed, pack_circuits
    from damavand_amd import workloads as wl
    circuits = wl.batch_workload(count, 1004)
    n_params = sum(wl.count_parameters(g) for _, g, _, _ in circuits)
    out = {"workload": f"{count} RY/RZ + CNOT-ring circuits, qubits U[12,18], layers U[4,12], Z/ZZ chain H",
           "circuits": count, "parameters": n_params}
    packed = pack_circuits(circuits)  # input preparation: flat arrays of the C ABI
    warm = pack_circuits(circuits[:8])
    res = {}
    for dtype in ("complex128", "complex64"):
        batch_forward_adjoint_packed(warm, dtype=dtype)  # warm-up
        N.profile_reset()
        N.profile_enable(True)
        t0 = time.perf_counter()
        e, g = batch_forward_adjoint_packed(packed, dtype=dtype)
        t = time.perf_counter() - t0
        N.profile_enable(False)
        launches, ms, nbytes = N.profile_get(N.KC_BATCH)
        res[dtype] = (e, g)
        key = "c128" if dtype == "complex128" else "c64"
        out[key] = {"call_s": t, "kernel_ms": ms, "circuits_per_s": count / (ms / 1e3),
                    "gradients_per_s": n_params / (ms / 1e3), "kernel_GBps": nbytes / (ms * 1e6)}
        log(f"[bench] batch {dtype}: kernel {ms:.1f} ms, call {t:.2f} s, {count / (ms / 1e3):.0f} circuits/s")
    e128, g128 = res["complex128"]
    e64, g64 = res["complex64"]
    out["c64_max_abs_err_energy_vs_c128"] = float(np.max(np.abs(e64 - e128)))
    out["c64_max_abs_err_grad_vs_c128"] = float(np.max(np.abs(g64 - g128)))
    return out


def bench_stress(n, layers, samples):
    """33-qubit complex128 (128 GiB): U3 on every qubit + 16 CNOT/CZ per layer."""
    from damavand_amd import _native as N
    from damavand_amd import workloads as wl
    gates = wl.stress_circuit(n, layers, 1005)
    c = build_circuit(n, gates)
    passes, phases, ops, hbm = c.plan_stats()
    c.forward()  # plans + warm-up
    c.reset_amplitudes()
    sync()
    N.profile_reset()
    N.profile_enable(True)
    t0 = time.perf_counter()
    c.forward()
    sync()
    t = time.perf_counter() - t0
    N.profile_enable(False)
    launches, ms, nbytes = N.profile_get(N.KC_TILE)
    t1 = time.perf_counter()
    ez = c.expectation_z()
    t_z = time.perf_counter() - t1
    t1 = time.perf_counter()
    s = c.sample(samples, seed=1005)
    t_s = time.perf_counter() - t1
    norm = c.state.norm_sqr()
    log(f"[bench] stress {n}q: {len(gates)} gates, {passes} passes, {t:.2f} s, {hbm / t / 1e9:.0f} GB/s; "
        f"<Z> {t_z * 1e3:.0f} ms, {samples} samples {t_s * 1e3:.0f} ms, |norm-1| {abs(norm - 1):.2e}")
    return {"workload": f"{n}-qubit complex128, {layers} layers of U3 on every qubit + 16 CNOT/CZ on seeded pairs",
            "gates": len(gates), "passes": passes, "forward_s": t, "GBps": hbm / t / 1e9,
            "tile_kernel_GBps": nbytes / (ms * 1e6) if ms else None, "expect_z_ms": t_z * 1e3,
            "samples": samples, "sample_ms": t_s * 1e3, "norm_err": abs(norm - 1), "z_first": ez[:4].tolist(),
            "distinct_samples": int(np.unique(s).size)}


def bench_cpu_baseline(n, layers, sample_gates):
    """Oracle (port of the reference's multithreading path) on this host."""
    import oracle
    from damavand_amd import workloads as wl
    gates = wl.layered_circuit(n, layers, 1002)[:sample_gates]
    t0 = time.perf_counter()
    st = oracle.OracleState(n)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    for kind, ctl, tgt, p in gates:
        st.apply_gate(kind, tgt, ctl, p)
    t = time.perf_counter() - t0
    algo = sum(wl.algorithmic_bytes_per_gate(g, n) for g in gates)
    cores = oracle.num_threads()
    # one more gate on a single thread (the reference's per-gate pass without the thread pool)
    oracle.set_num_threads(1)
    g1 = wl.layered_circuit(n, layers, 1002)[sample_gates]
    t0 = time.perf_counter()
    st.apply_gate(g1[0], g1[2], g1[1], g1[3])
    t1 = time.perf_counter() - t0
    oracle.set_num_threads(cores)
    # the reference sampler (utils.rs:258-277): one (N+1)-long cumulative rebuilt per draw
    probs = st.probabilities()
    t0 = time.perf_counter()
    oracle.sample_uniforms(probs, np.array([0.25, 0.75]), rebuild=True)
    t_sample = (time.perf_counter() - t0) / 2
    del st, probs
    return {"value": algo / t / 1e9, "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"first {len(gates)} gates (H on qubits 0..{len(gates) - 1}) of the {n}-qubit layered "
                      f"circuit, complex128, {t:.1f} s ({t / len(gates):.2f} s/gate; state init {t_init:.1f} s)",
            "value_1_thread": wl.algorithmic_bytes_per_gate(g1, n) / t1 / 1e9,
            "sample_1_thread": f"gate {sample_gates} of the same circuit on one thread, {t1:.1f} s",
            "sampler_s_per_draw": t_sample,
            "sampler_note": f"reference sampler at {n} qubits, cumulative rebuilt per draw, mean of 2 draws"}


def measured_traffic(qubits, layers):
    """HBM bytes per fused-pass launch from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    factor the same counter reads on the copy kernel), or None when the
    summary does not cover this workload."""
    path = os.path.join(ROOT, "profiles", "k_tile_30q_layered_c128_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("qubits") != qubits or t.get("layers") != layers:
        return None
    return (t["fetch_kib_per_dispatch"] * 1024.0 * t["fetch_factor"] +
            t["write_kib_per_dispatch"] * 1024.0 * t["write_factor"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=5)
    ap.add_argument("--adjoint-qubits", type=int, default=24)
    ap.add_argument("--adjoint-layers", type=int, default=20)
    ap.add_argument("--shift-params", type=int, default=96)
    ap.add_argument("--cpu-gates", type=int, default=6)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-per-gate", action="store_true")
    ap.add_argument("--batch", type=int, default=0, help="also run the batched workload with this many circuits")
    ap.add_argument("--stress", type=int, default=0, help="also run the stress workload at this many qubits")
    ap.add_argument("--stress-layers", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the state lives on a single MI355X and this build has no "
              "multi-GPU path. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from damavand_amd import _native as N
    log(f"[bench] {N.version()}, devices {N.device_count()}")
    copy_gbps = bench_copy(16 << 30)  # a state-sized copy: 16 GiB read + 16 GiB written
    log(f"[bench] streaming copy ceiling {copy_gbps:.0f} GB/s")
    fw, circ, gates = bench_forward(args.qubits, args.layers, args.steps, args.warmup)
    per_gate = None if args.skip_per_gate else bench_per_gate(circ)
    if per_gate:
        log(f"[bench] per-gate mode: {per_gate['ms']:.0f} ms, kernels {per_gate['GBps']:.0f} GB/s")
    del circ
    adj = None if args.skip_adjoint else bench_adjoint(args.adjoint_qubits, args.adjoint_layers, args.shift_params)
    batch = bench_batch(args.batch) if args.batch else None
    stress = bench_stress(args.stress, args.stress_layers, 100000) if args.stress else None
    cpu = None if args.skip_cpu else bench_cpu_baseline(args.qubits, args.layers, args.cpu_gates)
    if cpu:
        log(f"[bench] cpu baseline {cpu['value']:.2f} GB/s on {cpu['cores']} threads")

    value = fw["hbm_bytes"] / (fw["ms_step"] * 1e6)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fw["ms_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"{args.qubits}-qubit complex128 layered circuit, {args.layers} layers "
                        f"({fw['gates']} gates: H, RX, RY, RZ on every qubit, CNOT ladder), one fused forward per step",
            "qubits": args.qubits, "layers": args.layers, "gates": fw["gates"], "fused_passes": fw["passes"],
            "state_bytes": 16 * 2 ** args.qubits, "state_pieces": fw["state_pieces"], "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(fw["achieved"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(fw["achieved"] / HBM_PEAK_GBPS, 4), "traffic": measured_traffic(args.qubits, args.layers),
            "kernel": "k_tile (fused tile pass)", "launches": fw["launches"],
            "avg_launch_ms": round(fw["avg_launch_ms"], 4), "bytes_per_launch": fw["per_launch_bytes"],
        },
        "cpu_baseline": cpu,
        "gates_per_s": round(fw["gates"] / (fw["ms_step"] / 1e3), 1),
        "copy_ceiling_GBps": round(copy_gbps, 1),
        "per_gate_mode": per_gate,
        "adjoint": adj,
    }
    if batch:
        out["batch"] = batch
    if stress:
        out["stress"] = stress
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
