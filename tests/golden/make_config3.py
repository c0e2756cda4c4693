"""Writes tests/golden/config3_oracle.npz: the oracle's energy, <Z_i> and
parameter-shift gradient entries of the 24-qubit, 960-parameter adjoint
workload (20 layers of RY/RZ on every qubit plus a CNOT ring, Z/ZZ chain
observable with seeded normal coefficients; damavand_amd/workloads.py,
ansatz_circuit(24, 20, 1003)).

    python tests/golden/make_config3.py [--params 16] [--threads 8]

The reference has no gradient code (src/qubit_backend/circuit.rs:529-546 is a
commented-out parameter shift), so its gradient is pinned by the oracle's
parameter shift, exact for e^{-i theta P / 2} rotations:
    dE/dtheta_k = (E(theta_k + pi/2) - E(theta_k - pi/2)) / 2.
Each shifted energy is a full oracle forward (circuit_multithreading.rs:9-54
restated in oracle/oracle.c).  The forward is run once, keeping the state in
front of every selected parameter's gate, so each shift replays only the
gates after it.  About 20 minutes on 8 cores; ~5 GiB of RAM.

Contents: energy, expect_z[24], idx[P] (parameter indices, seeded),
grad_shift[P], plus the circuit's identity (n, layers, seed, gate count).
"""
import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from damavand_amd import workloads as wl  # noqa: E402

N, LAYERS, SEED = 24, 20, 1003


def selected_parameters(n_params, count, seed):
    """Seeded parameter subset: sorted, distinct, from splitmix64."""
    rng = wl.SplitMix64(seed)
    chosen = set()
    while len(chosen) < count:
        chosen.add(rng.next() % n_params)
    return np.array(sorted(chosen), dtype=np.int64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--params", type=int, default=16)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.threads:
        oracle.set_num_threads(args.threads)
    gates, a, b = wl.ansatz_circuit(N, LAYERS, SEED)
    pairs = wl.chain_pairs(N)
    n_params = wl.count_parameters(gates)
    idx = selected_parameters(n_params, args.params, SEED + 7)
    # gate index of every parameter, in insertion order (circuit.rs:308-322)
    pgate = [i for i, g in enumerate(gates) if g[0] in wl.PARAMETRIC]
    want = {pgate[k]: k for k in idx}
    t0 = time.time()
    st = oracle.OracleState(N)
    saved = {}
    for gi, (kind, ctl, tgt, prm) in enumerate(gates):
        if gi in want:
            saved[gi] = st.amp.copy()
        st.apply_gate(kind, tgt, ctl, prm)
    energy = st.expect_zz(a, pairs, b)
    ez = st.expect_z()
    print(f"forward done in {time.time() - t0:.0f} s, E = {energy!r}", flush=True)
    grad = np.zeros(idx.size)
    for j, k in enumerate(idx):
        gi = pgate[k]
        vals = []
        for s in (0.5 * math.pi, -0.5 * math.pi):
            st.amp[:] = saved[gi]
            kind, ctl, tgt, prm = gates[gi]
            st.apply_gate(kind, tgt, ctl, (prm[0] + s,))
            for kind2, ctl2, tgt2, prm2 in gates[gi + 1:]:
                st.apply_gate(kind2, tgt2, ctl2, prm2)
            vals.append(st.expect_zz(a, pairs, b))
        grad[j] = 0.5 * (vals[0] - vals[1])
        del saved[gi]
        print(f"param {k} (gate {gi}): {grad[j]!r}  [{time.time() - t0:.0f} s]", flush=True)
    path = os.path.join(HERE, "config3_oracle.npz")
    np.savez_compressed(path, n=N, layers=LAYERS, seed=SEED, gates=len(gates), energy=energy, expect_z=ez,
                        idx=idx, grad_shift=grad)
    print(path, os.path.getsize(path), "bytes", flush=True)


if __name__ == "__main__":
    main()
