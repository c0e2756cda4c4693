// Launchers of the HIP kernels (host side).  All launch on the library stream
// and return a DMV_* status; they check their own grid and buffer assumptions.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace dmv {

// One fused pass of a tile plan over a state of n qubits in layout L (every
// state launcher below takes the state's layout L).  half: the kernel with
// half a tile of LDS, for passes whose transposes are all split.
int launch_tile(int dtype, int rlog, int tlog, bool nontemporal, bool half, void *state, int n, StateLayout L,
                const DevPass *passes, const DevPhase *phases, const DevOp *ops, uint32_t pass_idx);
// Whether a half-LDS tile kernel exists for the shape.
bool tile_half_exists(int dtype, int rlog, int tlog);
// Reverse-sweep tile pass over psi and lambda; the IP ops of the pass write
// one partial per (IP op, tile, wave) into
// partials[(ip_pos * tiles + tile) * waves + wave].
int launch_tile_adjoint(int dtype, int rlog, int tlog, bool nt, void *psi, void *lam, int n, StateLayout L,
                        const DevPass *passes, const DevPhase *phases, const DevOp *ops,
                        uint32_t pass_idx, double *partials);
// Every op of a list applied to a whole state held in one workgroup's LDS.
int launch_small(int dtype, void *state, int n, const DevOp *ops, uint32_t n_ops);
// The same for the adjoint reverse sweep (both states in LDS); IP ops write
// partials[slot].
int launch_small_adjoint(int dtype, void *psi, void *lam, int n, const DevOp *ops, uint32_t n_ops,
                         double *partials);
// Batched persistent launch: one circuit per queue entry.  Circuits with
// n < k run with psi and lambda in LDS (ops in `sops`, global positions);
// larger ones run tile passes over per-workgroup scratch states.
// IP ops (gradient outputs) per reverse pass that the batched kernel keeps
// per-wave sums for in LDS; the host rejects plans with more.
constexpr uint32_t kBatchMaxIp = 128;
struct BatchCircuit {
    uint32_t n, small;
    uint32_t fwd_begin, fwd_end;  // pass range (tile path) or op range (small path)
    uint32_t adj_begin, adj_end;
    uint32_t coef;                // a_q at coef[coef + q], b_q at coef[coef + n + q]
    uint32_t grad;                // gradient of parameter k at grad[grad + k]
};
struct BatchLaunch {
    const BatchCircuit *circ;
    const uint32_t *order;
    uint32_t count;
    uint32_t *counter;
    const DevPass *passes;
    const DevPhase *phases;
    const DevOp *ops;
    const DevOp *sops;
    const double *coef;
    void *scratch;        // grid * 2 * slot_amps amplitudes
    uint64_t slot_amps;
    double *energy;
    double *grad;
    uint32_t grid;
    int rlog, tlog;       // tile shape: 2^rlog amplitudes per thread, 2^tlog threads
    bool nt;              // streaming loads and stores of the scratch states (default shape only)
};
// Whether the batched kernel is built for this tile shape.
bool batch_kernel_exists(int dtype, int rlog, int tlog);
int launch_batch(int dtype, const BatchLaunch &L);

// One op as its own pass over the amplitudes it changes (global qubits).
int launch_gate(int dtype, void *state, int n, StateLayout L, const DevOp &op, int pairs_per_thread, bool nontemporal);

// --- states, reductions, sampling ---
int launch_init_basis(int dtype, void *state, int n, StateLayout L, uint64_t basis);
int launch_copy(void *dst, const void *src, uint64_t bytes);
// Interleaved amplitudes [begin, begin+count) -> split re/im doubles (device).
int launch_split(int dtype, const void *state, StateLayout L, uint64_t begin, uint64_t count, double *re,
                 double *im);
int launch_join(int dtype, void *state, StateLayout L, uint64_t begin, uint64_t count, const double *re,
                const double *im);
int launch_probs(int dtype, const void *state, StateLayout L, uint64_t begin, uint64_t count, double *out);
// Partial sums of |a|^2 over blocks of 2^block_bits amplitudes.
int launch_block_norms(int dtype, const void *state, int n, StateLayout L, int block_bits, double *partials);
// Deterministic single-workgroup sum of `count` doubles into *out (device).
int launch_sum(const double *in, uint64_t count, double *out);
int launch_expect_z(int dtype, const void *state, int n, StateLayout L, double *partials, double *out);
int launch_expect_zz(int dtype, const void *state, int n, StateLayout L, const double *a, int n_pairs,
                     const int *pi, const int *pj, const double *b, double *partials, double *out);
// lambda = h * psi for the Z/ZZ observable, plus <H> into *out.
int launch_apply_obs(int dtype, const void *psi, void *lam, int n, StateLayout L, const double *a, int n_pairs,
                     const int *pi, const int *pj, const double *b, double *partials, double *out);
int launch_inner(int dtype, const void *a, const void *b, int n, StateLayout L, double *partials, double *out2);
// Diagonal observable over up to kMaxTableWires qubits, h(x) = table[bits of x
// at q[0..n)]: <H> into *out, and lambda = h * psi when lam is not null.
constexpr int kMaxTableWires = 12;
struct TableWires {
    uint32_t n;
    uint8_t q[kMaxTableWires];
};
int launch_obs_table(int dtype, const void *psi, void *lam, int n, StateLayout L, const double *table,
                     const TableWires &W, double *partials, double *out);
// Row sums: out[slot_map ? slot_map[b] : b] = sum_t partials[b * n_tiles + t]
// for b < n_rows.
int launch_slot_sums(const double *partials, int n_rows, uint64_t n_tiles, const int *slot_map, double *out);
int launch_sample(int dtype, const void *state, int n, StateLayout L, int64_t k, uint64_t seed, const double *u,
                  double *block_cum, double *partials, int64_t *out);

}  // namespace dmv
