// Host-side circuit lowering and the fused-pass scheduler (no HIP calls).
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace dmv {

// A gate as added through the Circuit API (src/qubit_backend/circuit.rs:608-691).
struct Gate {
    int kind;
    int control;  // -1 for none
    int target;
    double p[3];
    bool observable;
};

// An operation on global qubits, before tiling.
struct HostOp {
    uint32_t kind;
    int target;
    uint64_t ctrl;
    double m[8];
    int slot;  // parameter index of an IP op, else -1
};

// Process-wide tunables (defined by the runtime; environment DMV_<NAME> or dmv_tune_set).
int64_t tune_value(const char *name, int64_t dflt);

bool gate_is_parametric(int kind);
int gate_num_params(int kind);
// Row-major 2x2 of a gate, as m00r m00i m01r m01i m10r m10i m11r m11i.
int gate_matrix(int kind, const double *p, double m[8]);
// Operation class of a gate kind (OP_GEN, OP_DIAG or OP_SWAP).
uint32_t gate_class(int kind);

// Non-observable gates in insertion order, one op per gate (circuit.rs:346-349).
std::vector<HostOp> forward_ops(const std::vector<Gate> &gates);
// Reverse sweep of the adjoint method: for gates j = G..1, an IP op for a
// parametrised gate (slot = its parameter index) followed by G_j^dagger.
std::vector<HostOp> adjoint_ops(const std::vector<Gate> &gates);
// Folds runs of uncontrolled one-qubit ops on the same qubit into one op
// (their matrix product).  Valid because every op in between acts on other
// qubits; IP ops and controlled ops end a run.
std::vector<HostOp> merge_ops(const std::vector<HostOp> &ops);

struct PlanParams {
    int n;          // qubits of the state
    int k;          // tile qubits (= rlog + tlog)
    int rlog;       // log2 amplitudes per thread
    int tlog;       // log2 threads per workgroup
    int low_bits;   // lowest qubits kept in lanes for coalescing
    int swz_bits;   // LDS swizzle group (3 for 16-byte, 4 for 8-byte amplitudes)
    int max_gen_per_pass;  // compute budget of a pass, in general 2x2 ops
    int nontemporal = 0;   // forward passes: streaming (nontemporal) loads and stores
    int half_lds = 0;      // forward passes: split transposes where a pass allows (half-LDS kernels exist)
    int gen_auto = 0;      // forward plans: build_plan picks max_gen_per_pass (8 or 12) by plan cost
    StateLayout layout = kPlainLayout;  // element map of the states the plan runs on
};

struct Plan {
    int max_gen = 0;             // the GEN budget per pass the plan was built with
    std::vector<DevPass> passes;
    std::vector<DevPhase> phases;
    std::vector<DevOp> ops;      // tile-local ops, grouped by phase
    std::vector<int> op_pass;    // pass of each op
    int n_slots = 0;             // 1 + largest IP slot
    // IP ops per pass: pass p holds ip_begin[p+1] - ip_begin[p] of them, whose
    // slots (parameter indices) are ip_slots[ip_begin[p] ..] in ip_pos order
    std::vector<uint32_t> ip_begin;
    std::vector<int> ip_slots;
    // Every op in execution order with global qubits, including the X/CNOT
    // ops that only relabel registers and emit no device op.
    std::vector<HostOp> logical;
    std::vector<int> logical_pass;
};

// Greedy scheduler: packs ops into passes of <= k qubits (always including the
// `low_bits` lowest qubits) so that every op commutes past the ops it skips,
// then splits each pass into register phases.
int build_plan(const std::vector<HostOp> &ops, const PlanParams &pp, Plan *out);
// Checks every invariant the tile kernels rely on (index ranges, layouts,
// register bits).  Called before a plan is uploaded.
int validate_plan(const Plan &plan, const PlanParams &pp);
// The plan's ops with global targets/controls, in execution order.
std::vector<HostOp> plan_global_ops(const Plan &plan, std::vector<int> *pass_of_op);

// Whole-state ops with global qubit positions (for the single-workgroup kernel).
std::vector<DevOp> global_dev_ops(const std::vector<HostOp> &ops);

}  // namespace dmv
