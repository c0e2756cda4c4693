// Library runtime: device context, states, circuits and their cached plans.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "common.hpp"
#include "plan.hpp"

namespace dmv {

// A launch-shape choice: the autotuner's result for one circuit, or an entry
// of the process-wide table keyed by (GPU architecture, qubits, dtype, kind).
// Fields <= 0 (nt < 0) are unset.
struct TuneChoice {
    int k = 0;        // tile qubits
    int tlog = 0;     // log2 threads per workgroup
    int max_gen = 0;  // general 2x2 ops per fused pass
    int nt = -1;      // forward passes: nontemporal loads and stores (0/1)
    int wgs = 0;      // batched kernel: workgroups per CU
    double ms = 0.0;  // measured time of the choice
    bool set() const { return k > 0; }
};

// A device buffer made of kSpreadPieces physical allocations mapped back to
// back into one reserved address range (spread_alloc).
constexpr int kSpreadPieces = 8;
struct SpreadBuffer {
    void *va = nullptr;
    uint64_t bytes = 0;
    uint64_t piece = 0;
    hipMemGenericAllocationHandle_t h[kSpreadPieces] = {};
};

}  // namespace dmv

struct dmv_state {
    int n = 0;
    int dtype = DMV_C128;
    void *amp = nullptr;          // 2^n interleaved amplitudes, element order `layout`
    uint64_t bytes = 0;           // 2^n amplitudes' worth (algorithmic size)
    uint64_t alloc_bytes = 0;     // allocation size
    dmv::SpreadBuffer spread;     // set when amp is a spread buffer (else amp came from alloc)
    dmv::StateLayout layout{};    // element order of amp, fixed at creation
    double *ws = nullptr;         // reduction partials
    uint64_t ws_doubles = 0;
    double *cum = nullptr;        // sampler block prefix
    double *scalars = nullptr;    // small device outputs (64 doubles)
    double *stage = nullptr;      // split re/im staging of getters and setters (on first use)
    uint64_t stage_cap = 0;       // amplitudes per staged chunk
    double *terms = nullptr;      // Z/ZZ observable terms of dmv_expect_zz (on first use)
    uint64_t cap_terms = 0;
    void *bounce = nullptr;       // pinned host buffer (two halves) of device-to-host copies (on first use)
    hipEvent_t bounce_ev[2] = {nullptr, nullptr};
};

namespace dmv {

// A fused plan uploaded to the device.
struct DevicePlan {
    bool valid = false;
    uint64_t version = ~uint64_t(0);
    uint64_t build_id = 0;        // bumped on every rebuild
    uint64_t epoch = 0;           // tuning epoch the plan was built under
    int dtype = -1;
    bool small = false;           // whole-state kernel instead of tile passes
    PlanParams pp{};
    Plan plan;
    std::vector<DevOp> small_ops;
    DevPass *d_passes = nullptr;
    DevPhase *d_phases = nullptr;
    DevOp *d_ops = nullptr;
    int *d_ip_slots = nullptr;    // plan.ip_slots (reverse-sweep plans)
    size_t cap_passes = 0, cap_phases = 0, cap_ops = 0, cap_ip = 0;
    // hipGraph of the forward passes on one state buffer
    hipGraphExec_t graph = nullptr;
    const void *graph_state = nullptr;
    uint64_t graph_build = ~uint64_t(0);
};

}  // namespace dmv

struct dmv_circuit {
    int n = 0;
    std::vector<dmv::Gate> gates;
    uint64_t version = 0;
    dmv::DevicePlan fwd, adj;
    dmv::TuneChoice tune_fwd, tune_adj;  // set by the autotuner for this gate list
    double *d_partials = nullptr;   // adjoint partials of one pass: per (IP op, tile, wave)
    uint64_t cap_partials = 0;
    double *d_terms = nullptr;      // observable terms
    uint64_t cap_terms = 0;
    double *d_grad = nullptr;
    uint64_t cap_grad = 0;
};

namespace dmv {

int ensure_device();              // initialises the context, selects the device
int synchronize();
int alloc(void **p, uint64_t bytes);
void release(void *p);
// Large state buffers whose physical pieces lie kSpreadStride apart (see
// spread_alloc.cpp); DMV_ERR_MEMORY when the device cannot do it, with no
// error left set, so that the caller falls back to alloc.
int spread_alloc(uint64_t bytes, SpreadBuffer *out);
void spread_free(SpreadBuffer *b);
int copy_h2d(void *dst, const void *src, uint64_t bytes);
int copy_d2h(void *dst, const void *src, uint64_t bytes);
int64_t tune_value(const char *name, int64_t dflt);
uint64_t tune_epoch();

// Layout a new n-qubit state gets (tunables page_fold_span_bits, page_fold).
StateLayout default_layout(int n, int dtype);
// Tile shape for a forward (adjoint = false) or reverse sweep on n qubits
// over states of the given layout.  Each field comes from the first of: the
// circuit's own choice `pin` (may be null), the tunables set by
// dmv_tune_set or DMV_<NAME>, the tuning table, the defaults.
PlanParams tile_params(int n, int dtype, bool adjoint, StateLayout layout, const TuneChoice *pin = nullptr);
// Builds (if stale) and uploads the plan for `ops` into `dp`, for states of
// the given layout.
int prepare_plan(DevicePlan &dp, uint64_t version, int n, int dtype, bool adjoint,
                 const std::vector<HostOp> &ops, StateLayout layout, const TuneChoice *pin = nullptr);
// Runs the passes of a prepared forward plan on s.
int run_forward_plan(DevicePlan &dp, dmv_state *s);

// The tuning table: "<arch>/n<qubits>/<c128|c64>/<forward|adjoint|batch>".
// On first use it is filled from the JSON-lines file named by DMV_TUNE_CACHE
// (later lines win); tune_store adds to the table and appends to that file.
std::string tune_key(int n, int dtype, const char *kind);
bool tune_lookup(const std::string &key, TuneChoice *out);
void tune_store(const std::string &key, const TuneChoice &c);
void free_plan(DevicePlan &dp);

int state_create(int n, int dtype, dmv_state **out);
void state_destroy(dmv_state *s);
int state_reset(dmv_state *s, uint64_t basis, bool zero);
int state_get(dmv_state *s, double *re, double *im);
int state_set(dmv_state *s, const double *re, const double *im);
int state_probs(dmv_state *s, double *host);

int circuit_forward(dmv_circuit *c, dmv_state *s, int mode);
// The observable of an adjoint gradient: the Z/ZZ form
// H = sum_q a_q Z_q + sum_t b_t Z_pi[t] Z_pj[t] when table is null, else the
// diagonal H|x> = table[sum_j bit_{wires[j]}(x) 2^j] |x>.
struct Observable {
    const double *a = nullptr;
    int n_pairs = 0;
    const int *pi = nullptr, *pj = nullptr;
    const double *b = nullptr;
    int n_wires = 0;
    const int *wires = nullptr;
    const double *table = nullptr;
};
// Adjoint gradient (dmv_circuit_adjoint) with the reverse-sweep plan `adj`
// built under `pin`.
int circuit_adjoint(dmv_circuit *c, dmv_state *psi, dmv_state *lam, const Observable &obs, double *energy,
                    double *grad, DevicePlan &adj, const TuneChoice *pin);

// k samples of s into the device array d_out (uniforms d_u, or the seeded
// stream when d_u is null), using only the state's workspace: no allocation.
int sample_to_device(dmv_state *s, int64_t k, uint64_t seed, const double *d_u, int64_t *d_out);

}  // namespace dmv
