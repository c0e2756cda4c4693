// The dmv_* C ABI of libdamavand-gpu.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace dmv {
int profile_enable(int on);
int profile_reset();
int profile_get(int kc, int64_t *launches, double *ms, double *bytes);
void tune_set(const char *name, int64_t v);
int batch_forward_adjoint(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                          const int *gate_control, const int *gate_target, const double *gate_param,
                          const int *coef_offset, const double *coef, const int *param_offset, double *energy,
                          double *grad, const TuneChoice *pin, double *kernel_ms);
int autotune_forward(dmv_circuit *c, dmv_state *s, int *chosen);
int autotune_adjoint(dmv_circuit *c, int dtype, int *chosen);
void tune_table_reset();
int autotune_batch(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                   const int *gate_control, const int *gate_target, const double *gate_param, const int *coef_offset,
                   const double *coef, const int *param_offset, int *chosen);
}  // namespace dmv

namespace dmv {
// k samples of s into the device array d_out (uniforms from d_u, or from the
// seeded stream when d_u is null).  Uses only the state's own workspace
// (block masses and their prefix) and allocates nothing.
int sample_to_device(dmv_state *s, int64_t k, uint64_t seed, const double *d_u, int64_t *d_out)
{
    // the reference panics when it cannot sample (circuit.rs:478-482): refuse a zero state
    const int bb = s->n < 12 ? s->n : 12;
    DMV_TRY(launch_block_norms(s->dtype, s->amp, s->n, s->layout, bb, s->ws));
    DMV_TRY(launch_sum(s->ws, uint64_t(1) << (s->n - bb), s->scalars));
    double total = 0.0;
    DMV_TRY(copy_d2h(&total, s->scalars, sizeof(double)));
    if (!(total > 0.0 && std::isfinite(total))) {
        set_error("could not sample from output distribution (total probability %g)", total);
        return DMV_ERR_STATE;
    }
    prof_begin(DMV_KC_SAMPLE);
    DMV_TRY(launch_sample(s->dtype, s->amp, s->n, s->layout, k, seed, d_u, s->cum, s->ws, d_out));
    prof_end(DMV_KC_SAMPLE, (double)s->bytes);
    return DMV_OK;
}
}  // namespace dmv

using namespace dmv;

#define CHECK_STATE(s) DMV_CHECK_ARG((s) != nullptr && (s)->amp != nullptr, "null state")
#define CHECK_CIRCUIT(c) DMV_CHECK_ARG((c) != nullptr, "null circuit")

// Uploads a Z/ZZ observable into a device buffer laid out as
// a[n] (double) | b[n_pairs] (double) | pi[n_pairs] (int) | pj[n_pairs] (int).
static int upload_terms(double **buf, uint64_t *cap, int n, const double *a, int n_pairs, const int *pi,
                        const int *pj, const double *b, const double **da, const double **db, const int **dpi,
                        const int **dpj)
{
    DMV_CHECK_ARG(n_pairs >= 0 && n_pairs <= 128, "at most 128 ZZ terms");
    DMV_CHECK_ARG(n_pairs == 0 || (pi && pj && b), "null ZZ term arrays");
    for (int t = 0; t < n_pairs; t++)
        DMV_CHECK_ARG(pi[t] >= 0 && pi[t] < n && pj[t] >= 0 && pj[t] < n, "ZZ term qubit out of range");
    const uint64_t need = (uint64_t)n + (uint64_t)n_pairs * 2 + 8;
    if (need > *cap) {
        release(*buf);
        *buf = nullptr;
        DMV_TRY(alloc((void **)buf, need * sizeof(double)));
        *cap = need;
    }
    std::vector<double> host(need, 0.0);
    for (int q = 0; q < n; q++)
        host[q] = a ? a[q] : 0.0;
    for (int t = 0; t < n_pairs; t++)
        host[n + t] = b[t];
    int *ip = reinterpret_cast<int *>(host.data() + n + n_pairs);
    for (int t = 0; t < n_pairs; t++) {
        ip[t] = pi[t];
        ip[n_pairs + t] = pj[t];
    }
    DMV_TRY(copy_h2d(*buf, host.data(), need * sizeof(double)));
    *da = *buf;
    *db = *buf + n;
    *dpi = reinterpret_cast<const int *>(*buf + n + n_pairs);
    *dpj = *dpi + n_pairs;
    return DMV_OK;
}

// Uploads the table of a diagonal observable (2^n_wires doubles) and checks
// its wires: distinct, in [0, n).
static int upload_table(double **buf, uint64_t *cap, int n, int n_wires, const int *wires, const double *table,
                        TableWires *W)
{
    DMV_CHECK_ARG(n_wires >= 0 && n_wires <= kMaxTableWires, "a table observable has 0 to %d wires", kMaxTableWires);
    DMV_CHECK_ARG(n_wires == 0 || wires, "null wire array");
    DMV_CHECK_ARG(table, "null table");
    W->n = (uint32_t)n_wires;
    uint64_t seen = 0;
    for (int j = 0; j < n_wires; j++) {
        DMV_CHECK_ARG(wires[j] >= 0 && wires[j] < n, "table observable wire %d out of range", wires[j]);
        DMV_CHECK_ARG(!((seen >> wires[j]) & 1), "table observable wire %d repeated", wires[j]);
        seen |= uint64_t(1) << wires[j];
        W->q[j] = (uint8_t)wires[j];
    }
    const uint64_t need = uint64_t(1) << n_wires;
    for (uint64_t i = 0; i < need; i++)
        DMV_CHECK_ARG(std::isfinite(table[i]), "non-finite table entry %llu", (unsigned long long)i);
    if (need > *cap) {
        release(*buf);
        *buf = nullptr;
        DMV_TRY(alloc((void **)buf, need * sizeof(double)));
        *cap = need;
    }
    return copy_h2d(*buf, table, need * sizeof(double));
}

int dmv::circuit_adjoint(dmv_circuit *c, dmv_state *psi, dmv_state *lam, const Observable &obs, double *energy,
                         double *grad, DevicePlan &dp, const TuneChoice *pin)
{
    CHECK_CIRCUIT(c);
    CHECK_STATE(psi);
    CHECK_STATE(lam);
    DMV_CHECK_ARG(psi != lam, "psi and lambda must be different states");
    DMV_CHECK_ARG(psi->n == c->n && lam->n == c->n && psi->dtype == lam->dtype, "state/circuit size mismatch");
    DMV_CHECK_ARG(same_layout(psi->layout, lam->layout), "psi and lambda differ in element order");
    DMV_TRY(ensure_device());
    int np = 0;
    dmv_circuit_num_parameters(c, &np);
    // psi = U|0>
    DMV_TRY(state_reset(psi, 0, false));
    DMV_TRY(circuit_forward(c, psi, DMV_FORWARD_FUSED));
    // lambda = H psi, E = <psi|H|psi>
    if (obs.table) {
        TableWires W;
        DMV_TRY(upload_table(&c->d_terms, &c->cap_terms, c->n, obs.n_wires, obs.wires, obs.table, &W));
        prof_begin(DMV_KC_REDUCE);
        DMV_TRY(launch_obs_table(psi->dtype, psi->amp, lam->amp, c->n, psi->layout, c->d_terms, W, psi->ws,
                                 psi->scalars));
        prof_end(DMV_KC_REDUCE, 3.0 * (double)psi->bytes);
    } else {
        const double *da, *db;
        const int *dpi, *dpj;
        DMV_TRY(upload_terms(&c->d_terms, &c->cap_terms, c->n, obs.a, obs.n_pairs, obs.pi, obs.pj, obs.b, &da, &db,
                             &dpi, &dpj));
        prof_begin(DMV_KC_REDUCE);
        DMV_TRY(launch_apply_obs(psi->dtype, psi->amp, lam->amp, c->n, psi->layout, da, obs.n_pairs, dpi, dpj, db,
                                 psi->ws, psi->scalars));
        prof_end(DMV_KC_REDUCE, 3.0 * (double)psi->bytes);
    }
    // reverse sweep
    DMV_TRY(prepare_plan(dp, c->version, c->n, psi->dtype, true, adjoint_ops(c->gates), psi->layout, pin));
    if (dp.plan.n_slots != np) {
        set_error("internal: adjoint plan has %d slots for %d parameters", dp.plan.n_slots, np);
        return DMV_ERR_STATE;
    }
    // gradient partials: per (IP op, tile, wave) on the tile path, per slot
    // on the whole-state path.  The tile path keeps the rows of as many
    // consecutive passes as fit in kPartialBudget bytes (at least one pass)
    // and sums them into grad with one launch whenever the next pass would
    // not fit, so the buffer stays bounded at any state size.
    constexpr uint64_t kPartialBudget = uint64_t(256) << 20;
    uint64_t per_ip = 1, rows = (uint64_t)(np > 0 ? np : 1);
    if (!dp.small) {
        per_ip = (uint64_t(1) << (c->n - dp.pp.k)) * ((uint64_t(1) << dp.pp.tlog) / 64);
        uint64_t widest = 1;
        for (size_t p = 0; p < dp.plan.passes.size(); p++)
            widest = std::max<uint64_t>(widest, dp.plan.ip_begin[p + 1] - dp.plan.ip_begin[p]);
        rows = std::max(widest, std::min<uint64_t>(std::max<uint64_t>(1, dp.plan.ip_slots.size()),
                                                   kPartialBudget / (per_ip * sizeof(double))));
    }
    if (rows * per_ip > c->cap_partials) {
        release(c->d_partials);
        c->d_partials = nullptr;
        DMV_TRY(alloc((void **)&c->d_partials, rows * per_ip * sizeof(double)));
        c->cap_partials = rows * per_ip;
    }
    if ((uint64_t)np + 1 > c->cap_grad) {
        release(c->d_grad);
        c->d_grad = nullptr;
        DMV_TRY(alloc((void **)&c->d_grad, ((uint64_t)np + 1) * sizeof(double)));
        c->cap_grad = (uint64_t)np + 1;
    }
    if (dp.small) {
        prof_begin(DMV_KC_ADJOINT);
        DMV_TRY(launch_small_adjoint(psi->dtype, psi->amp, lam->amp, c->n, dp.d_ops, (uint32_t)dp.small_ops.size(),
                                     c->d_partials));
        prof_end(DMV_KC_ADJOINT, 4.0 * (double)psi->bytes);
        DMV_TRY(launch_slot_sums(c->d_partials, np, 1, nullptr, c->d_grad));
    } else {
        const uint32_t n_pass = (uint32_t)dp.plan.passes.size();
        uint32_t pending = 0;  // first IP row not yet summed
        for (uint32_t p = 0; p < n_pass; p++) {
            const uint64_t row0 = dp.plan.ip_begin[p] - pending;
            prof_begin(DMV_KC_ADJOINT);
            DMV_TRY(launch_tile_adjoint(psi->dtype, dp.pp.rlog, dp.pp.tlog, dp.pp.nontemporal != 0, psi->amp, lam->amp,
                                        c->n, psi->layout, dp.d_passes, dp.d_phases, dp.d_ops, p,
                                        c->d_partials + row0 * per_ip));
            prof_end(DMV_KC_ADJOINT, 4.0 * (double)psi->bytes);
            const uint32_t done = dp.plan.ip_begin[p + 1];
            if (p + 1 == n_pass || dp.plan.ip_begin[p + 2] - pending > rows) {
                DMV_TRY(launch_slot_sums(c->d_partials, (int)(done - pending), per_ip, dp.d_ip_slots + pending, c->d_grad));
                pending = done;
            }
        }
    }
    if (energy)
        DMV_TRY(copy_d2h(energy, psi->scalars, sizeof(double)));
    if (grad && np > 0)
        DMV_TRY(copy_d2h(grad, c->d_grad, sizeof(double) * np));
    return DMV_OK;
}

extern "C" {

const char *dmv_last_error(void) { return last_error(); }

const char *dmv_version(void) { return "damavand-mi355x 0.1 (gfx950)"; }

int dmv_device_count(int *count)
{
    DMV_CHECK_ARG(count != nullptr, "null pointer");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return DMV_OK;
}

int dmv_device_memory(uint64_t *free_bytes, uint64_t *total_bytes)
{
    DMV_TRY(ensure_device());
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) {
        set_error("hipMemGetInfo failed");
        return DMV_ERR_HIP;
    }
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return DMV_OK;
}

int dmv_state_bytes(int num_qubits, int dtype, uint64_t *bytes)
{
    DMV_CHECK_ARG(bytes != nullptr, "null pointer");
    DMV_CHECK_ARG(num_qubits >= 1 && num_qubits <= kMaxQubits, "num_qubits out of range");
    const uint64_t amp = (uint64_t(1) << num_qubits) * (uint64_t)amp_bytes(dtype);
    const int B = num_qubits < 12 ? num_qubits : 12;
    const uint64_t nblk = uint64_t(1) << (num_qubits - B);
    *bytes = amp + (nblk * 13 + 64) * 8 * 2 + 64 * 8;
    return DMV_OK;
}

int dmv_max_qubits(int dtype, int *max_qubits)
{
    DMV_CHECK_ARG(max_qubits != nullptr, "null pointer");
    uint64_t f = 0;
    DMV_TRY(dmv_device_memory(&f, nullptr));
    int best = 0;
    for (int q = 1; q <= kMaxQubits; q++) {
        uint64_t need = 0;
        dmv_state_bytes(q, dtype, &need);
        if (need + (uint64_t(1) << 28) <= f)  // keep 256 MiB of headroom for staging
            best = q;
    }
    *max_qubits = best;
    return DMV_OK;
}

// ------------------------------------------------------------------- states --
int dmv_state_create(int num_qubits, int dtype, dmv_state **out)
{
    if (num_qubits >= 1 && num_qubits <= kMaxQubits && (dtype == DMV_C128 || dtype == DMV_C64)) {
        DMV_TRY(ensure_device());
        uint64_t need = 0, f = 0;
        dmv_state_bytes(num_qubits, dtype, &need);
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
            f = fr;
            if (need > f) {
                set_error("a %d-qubit state needs %llu bytes of device memory, %llu are free", num_qubits,
                          (unsigned long long)need, (unsigned long long)f);
                return DMV_ERR_MEMORY;
            }
        }
    }
    return state_create(num_qubits, dtype, out);
}

int dmv_state_destroy(dmv_state *s)
{
    if (s)
        state_destroy(s);
    return DMV_OK;
}

int dmv_state_num_qubits(const dmv_state *s, int *num_qubits)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(num_qubits != nullptr, "null pointer");
    *num_qubits = s->n;
    return DMV_OK;
}

int dmv_state_reset(dmv_state *s, uint64_t basis_index)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(basis_index < (uint64_t(1) << s->n), "basis index out of range");
    return state_reset(s, basis_index, false);
}

int dmv_state_get(dmv_state *s, double *re, double *im)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(re != nullptr || im != nullptr, "null output arrays");
    return state_get(s, re, im);
}

int dmv_state_set(dmv_state *s, const double *re, const double *im)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(re != nullptr && im != nullptr, "null input arrays");
    return state_set(s, re, im);
}

int dmv_state_copy(dmv_state *dst, const dmv_state *src)
{
    CHECK_STATE(dst);
    CHECK_STATE(src);
    DMV_CHECK_ARG(dst->n == src->n && dst->dtype == src->dtype, "states differ in size or dtype");
    DMV_CHECK_ARG(same_layout(dst->layout, src->layout), "states differ in element order (created under other settings)");
    DMV_TRY(ensure_device());
    prof_begin(DMV_KC_COPY);
    DMV_TRY(launch_copy(dst->amp, src->amp, dst->alloc_bytes));
    prof_end(DMV_KC_COPY, 2.0 * (double)dst->bytes);
    return DMV_OK;
}

int dmv_state_device_ptr(dmv_state *s, void **ptr, uint64_t *bytes)
{
    CHECK_STATE(s);
    if (ptr)
        *ptr = s->amp;
    if (bytes)
        *bytes = s->alloc_bytes;
    return DMV_OK;
}

int dmv_state_pieces(const dmv_state *s, int *pieces)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(pieces != nullptr, "null output pointer");
    *pieces = s->spread.va ? kSpreadPieces : 1;
    return DMV_OK;
}

int dmv_state_layout(const dmv_state *s, int *shift, int *block_bits, int *fold_bits)
{
    CHECK_STATE(s);
    const StateLayout L = s->layout;
    if (shift)
        *shift = (int)L.shift;
    if (block_bits)
        *block_bits = (int)L.blk;
    if (fold_bits)
        *fold_bits = (int)L.fold;
    return DMV_OK;
}

// ------------------------------------------------------------ single ops --
static int classify(const double *m_re, const double *m_im, DevOp *op)
{
    for (int i = 0; i < 4; i++) {
        DMV_CHECK_ARG(std::isfinite(m_re[i]) && std::isfinite(m_im[i]), "non-finite matrix entry");
        op->m[2 * i] = m_re[i];
        op->m[2 * i + 1] = m_im[i];
    }
    const bool offdiag_zero = m_re[1] == 0 && m_im[1] == 0 && m_re[2] == 0 && m_im[2] == 0;
    const bool is_x = m_re[0] == 0 && m_im[0] == 0 && m_re[3] == 0 && m_im[3] == 0 && m_re[1] == 1 &&
                      m_im[1] == 0 && m_re[2] == 1 && m_im[2] == 0;
    op->kind = offdiag_zero ? OP_DIAG : (is_x ? OP_SWAP : OP_GEN);
    return DMV_OK;
}

int dmv_apply_matrix(dmv_state *s, const double *m_re, const double *m_im, uint64_t control_mask, int target)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(m_re && m_im, "null matrix");
    DMV_CHECK_ARG(target >= 0 && target < s->n, "target qubit %d out of range for %d qubits", target, s->n);
    DMV_CHECK_ARG((control_mask >> s->n) == 0, "control qubit out of range");
    DMV_CHECK_ARG(!((control_mask >> target) & 1), "control and target coincide");
    DMV_TRY(ensure_device());
    DevOp op;
    std::memset(&op, 0, sizeof(op));
    DMV_TRY(classify(m_re, m_im, &op));
    op.target = (uint32_t)target;
    op.ctrl = control_mask;
    op.slot = -1;
    const uint64_t n_amp = uint64_t(1) << s->n;
    const bool phase_only = op.kind == OP_DIAG && op.m[0] == 1.0 && op.m[1] == 0.0;
    const double touched = (double)(n_amp >> (popcount64(control_mask) + (phase_only ? 1 : 0)));
    prof_begin(DMV_KC_GATE);
    DMV_TRY(launch_gate(s->dtype, s->amp, s->n, s->layout, op, gate_pairs_per_thread(), gate_nontemporal(op.target, s->dtype)));
    prof_end(DMV_KC_GATE, 2.0 * touched * amp_bytes(s->dtype));
    return DMV_OK;
}

int dmv_apply_gate(dmv_state *s, int kind, int control, int target, const double *params)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(control < s->n && control != target, "invalid control qubit %d", control);
    if (gate_num_params(kind) > 0)
        DMV_CHECK_ARG(params != nullptr, "gate needs parameters");
    double m[8];
    DMV_TRY(gate_matrix(kind, params, m));
    if ((kind == DMV_GATE_CNOT || kind == DMV_GATE_CZ))
        DMV_CHECK_ARG(control >= 0, "controlled gate without a control qubit");
    const double re[4] = {m[0], m[2], m[4], m[6]}, im[4] = {m[1], m[3], m[5], m[7]};
    return dmv_apply_matrix(s, re, im, control >= 0 ? (uint64_t(1) << control) : 0, target);
}

// ------------------------------------------------------------- measurement --
int dmv_probabilities(dmv_state *s, double *host_probabilities)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(host_probabilities != nullptr, "null output");
    return state_probs(s, host_probabilities);
}

int dmv_norm_sqr(dmv_state *s, double *out)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(out != nullptr, "null output");
    DMV_TRY(ensure_device());
    const int B = s->n < 12 ? s->n : 12;
    prof_begin(DMV_KC_REDUCE);
    DMV_TRY(launch_block_norms(s->dtype, s->amp, s->n, s->layout, B, s->ws));
    DMV_TRY(launch_sum(s->ws, uint64_t(1) << (s->n - B), s->scalars));
    prof_end(DMV_KC_REDUCE, (double)s->bytes);
    return copy_d2h(out, s->scalars, sizeof(double));
}

int dmv_expect_z(dmv_state *s, double *out)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(out != nullptr, "null output");
    DMV_TRY(ensure_device());
    prof_begin(DMV_KC_REDUCE);
    DMV_TRY(launch_expect_z(s->dtype, s->amp, s->n, s->layout, s->ws, s->scalars));
    prof_end(DMV_KC_REDUCE, (double)s->bytes);
    return copy_d2h(out, s->scalars, sizeof(double) * s->n);
}

int dmv_expect_zz(dmv_state *s, const double *a, int n_pairs, const int *pi, const int *pj, const double *b,
                  double *out)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(out != nullptr, "null output");
    DMV_TRY(ensure_device());
    const double *da, *db;
    const int *dpi, *dpj;
    DMV_TRY(upload_terms(&s->terms, &s->cap_terms, s->n, a, n_pairs, pi, pj, b, &da, &db, &dpi, &dpj));
    prof_begin(DMV_KC_REDUCE);
    DMV_TRY(launch_expect_zz(s->dtype, s->amp, s->n, s->layout, da, n_pairs, dpi, dpj, db, s->ws, s->scalars));
    prof_end(DMV_KC_REDUCE, (double)s->bytes);
    return copy_d2h(out, s->scalars, sizeof(double));
}

int dmv_expect_diag(dmv_state *s, int n_wires, const int *wires, const double *table, double *out)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(out != nullptr, "null output");
    DMV_TRY(ensure_device());
    TableWires W;
    DMV_TRY(upload_table(&s->terms, &s->cap_terms, s->n, n_wires, wires, table, &W));
    prof_begin(DMV_KC_REDUCE);
    DMV_TRY(launch_obs_table(s->dtype, s->amp, nullptr, s->n, s->layout, s->terms, W, s->ws, s->scalars));
    prof_end(DMV_KC_REDUCE, (double)s->bytes);
    return copy_d2h(out, s->scalars, sizeof(double));
}

int dmv_inner_product(dmv_state *a, dmv_state *b, double *out_re_im)
{
    CHECK_STATE(a);
    CHECK_STATE(b);
    DMV_CHECK_ARG(out_re_im != nullptr, "null output");
    DMV_CHECK_ARG(a->n == b->n && a->dtype == b->dtype, "states differ in size or dtype");
    DMV_CHECK_ARG(same_layout(a->layout, b->layout), "states differ in element order (created under other settings)");
    DMV_TRY(ensure_device());
    prof_begin(DMV_KC_REDUCE);
    DMV_TRY(launch_inner(a->dtype, a->amp, b->amp, a->n, a->layout, a->ws, a->scalars));
    prof_end(DMV_KC_REDUCE, 2.0 * (double)a->bytes);
    return copy_d2h(out_re_im, a->scalars, 2 * sizeof(double));
}

static int sample_impl(dmv_state *s, int64_t k, uint64_t seed, const double *u_host, int64_t *out)
{
    CHECK_STATE(s);
    DMV_CHECK_ARG(out != nullptr, "null output");
    DMV_CHECK_ARG(k >= 0, "negative sample count");
    if (k == 0)
        return DMV_OK;
    if (u_host)
        for (int64_t i = 0; i < k; i++)
            DMV_CHECK_ARG(u_host[i] > 0.0 && u_host[i] <= 1.0, "uniform %lld is not in (0, 1]", (long long)i);
    DMV_TRY(ensure_device());
    int64_t *d_out = nullptr;
    double *d_u = nullptr;
    DMV_TRY(alloc((void **)&d_out, (uint64_t)k * sizeof(int64_t)));
    int rc = DMV_OK;
    if (u_host) {
        rc = alloc((void **)&d_u, (uint64_t)k * sizeof(double));
        if (rc == DMV_OK)
            rc = copy_h2d(d_u, u_host, (uint64_t)k * sizeof(double));
    }
    if (rc == DMV_OK)
        rc = sample_to_device(s, k, seed, d_u, d_out);
    if (rc == DMV_OK)
        rc = copy_d2h(out, d_out, (uint64_t)k * sizeof(int64_t));
    release(d_out);
    release(d_u);
    return rc;
}

int dmv_sample(dmv_state *s, int64_t k, uint64_t seed, int64_t *out) { return sample_impl(s, k, seed, nullptr, out); }

int dmv_sample_uniforms(dmv_state *s, int64_t k, const double *u, int64_t *out)
{
    DMV_CHECK_ARG(u != nullptr || k == 0, "null uniforms");
    return sample_impl(s, k, 0, u, out);
}

// ---------------------------------------------------------------- circuits --
int dmv_circuit_create(int num_qubits, dmv_circuit **out)
{
    DMV_CHECK_ARG(out != nullptr, "null output pointer");
    DMV_CHECK_ARG(num_qubits >= 1 && num_qubits <= kMaxQubits, "num_qubits must be in [1, %d]", kMaxQubits);
    dmv_circuit *c = new dmv_circuit();
    c->n = num_qubits;
    *out = c;
    return DMV_OK;
}

int dmv_circuit_destroy(dmv_circuit *c)
{
    if (!c)
        return DMV_OK;
    if (c->fwd.valid || c->adj.valid || c->d_partials || c->d_terms || c->d_grad) {
        free_plan(c->fwd);
        free_plan(c->adj);
        release(c->d_partials);
        release(c->d_terms);
        release(c->d_grad);
    }
    delete c;
    return DMV_OK;
}

int dmv_circuit_add_gate(dmv_circuit *c, int kind, int control, int target, const double *params, int is_observable,
                         int *index)
{
    CHECK_CIRCUIT(c);
    DMV_CHECK_ARG(kind >= DMV_GATE_H && kind <= DMV_GATE_CZ, "unknown gate kind %d", kind);
    DMV_CHECK_ARG(target >= 0 && target < c->n, "target qubit %d out of range for %d qubits", target, c->n);
    const bool controlled = kind == DMV_GATE_CNOT || kind == DMV_GATE_CZ;
    if (controlled)
        DMV_CHECK_ARG(control >= 0 && control < c->n && control != target, "invalid control qubit %d", control);
    else
        DMV_CHECK_ARG(control == -1, "gate kind %d takes no control", kind);
    if (is_observable)
        DMV_CHECK_ARG(kind == DMV_GATE_X || kind == DMV_GATE_Y || kind == DMV_GATE_Z,
                      "only Pauli gates can be observables");
    Gate g;
    g.kind = kind;
    g.control = controlled ? control : -1;
    g.target = target;
    g.p[0] = g.p[1] = g.p[2] = 0.0;
    const int np = gate_num_params(kind);
    if (np > 0) {
        DMV_CHECK_ARG(params != nullptr, "gate kind %d needs %d parameter(s)", kind, np);
        for (int i = 0; i < np; i++) {
            DMV_CHECK_ARG(std::isfinite(params[i]), "non-finite gate parameter");
            g.p[i] = params[i];
        }
    }
    g.observable = is_observable != 0;
    c->gates.push_back(g);
    c->version++;
    if (index)
        *index = (int)c->gates.size() - 1;
    return DMV_OK;
}

int dmv_circuit_num_gates(const dmv_circuit *c, int *count)
{
    CHECK_CIRCUIT(c);
    DMV_CHECK_ARG(count != nullptr, "null pointer");
    *count = (int)c->gates.size();
    return DMV_OK;
}

int dmv_circuit_num_parameters(const dmv_circuit *c, int *count)
{
    CHECK_CIRCUIT(c);
    DMV_CHECK_ARG(count != nullptr, "null pointer");
    int k = 0;
    for (const Gate &g : c->gates)
        if (gate_is_parametric(g.kind))
            k++;
    *count = k;
    return DMV_OK;
}

int dmv_circuit_set_parameters(dmv_circuit *c, const double *p, int count)
{
    CHECK_CIRCUIT(c);
    DMV_CHECK_ARG(count >= 0 && (p != nullptr || count == 0), "bad parameter array");
    // every value is checked before any gate changes, so a rejected call
    // leaves the circuit (and its cached plans) as they were
    int np = 0;
    for (const Gate &g : c->gates)
        if (gate_is_parametric(g.kind))
            np++;
    const int used = count < np ? count : np;
    for (int k = 0; k < used; k++)
        DMV_CHECK_ARG(std::isfinite(p[k]), "non-finite parameter %d", k);
    int k = 0;
    for (Gate &g : c->gates) {
        if (!gate_is_parametric(g.kind))
            continue;
        if (k >= used)
            break;
        g.p[0] = p[k++];
    }
    c->version++;
    return DMV_OK;
}

int dmv_circuit_get_parameters(const dmv_circuit *c, double *p, int count)
{
    CHECK_CIRCUIT(c);
    int k = 0;
    for (const Gate &g : c->gates)
        if (gate_is_parametric(g.kind) && k < count)
            p[k++] = g.p[0];
    return DMV_OK;
}

int dmv_circuit_observables(const dmv_circuit *c, int *targets, int capacity, int *count)
{
    CHECK_CIRCUIT(c);
    int k = 0;
    for (const Gate &g : c->gates)
        if (g.observable) {
            if (targets && k < capacity)
                targets[k] = g.target;
            k++;
        }
    if (count)
        *count = k;
    return DMV_OK;
}

int dmv_circuit_clear(dmv_circuit *c)
{
    CHECK_CIRCUIT(c);
    c->gates.clear();
    c->version++;
    c->tune_fwd = TuneChoice();  // tuned for the gates just removed
    c->tune_adj = TuneChoice();
    return DMV_OK;
}

int dmv_circuit_forward(dmv_circuit *c, dmv_state *s, int mode)
{
    CHECK_CIRCUIT(c);
    CHECK_STATE(s);
    return circuit_forward(c, s, mode);
}

int dmv_circuit_adjoint(dmv_circuit *c, dmv_state *psi, dmv_state *lam, const double *a, int n_pairs, const int *pi,
                        const int *pj, const double *b, double *energy, double *grad)
{
    Observable obs;
    obs.a = a;
    obs.n_pairs = n_pairs;
    obs.pi = pi;
    obs.pj = pj;
    obs.b = b;
    return circuit_adjoint(c, psi, lam, obs, energy, grad, c->adj, &c->tune_adj);
}

int dmv_circuit_adjoint_diag(dmv_circuit *c, dmv_state *psi, dmv_state *lam, int n_wires, const int *wires,
                             const double *table, double *energy, double *grad)
{
    DMV_CHECK_ARG(table != nullptr, "null table");
    Observable obs;
    obs.n_wires = n_wires;
    obs.wires = wires;
    obs.table = table;
    return circuit_adjoint(c, psi, lam, obs, energy, grad, c->adj, &c->tune_adj);
}

int dmv_circuit_plan_stats(dmv_circuit *c, int dtype, int *passes, int *phases, int *ops, double *hbm_bytes)
{
    CHECK_CIRCUIT(c);
    const int small_cap = dtype == DMV_C64 ? kSmallMaxQubitsC64 : kSmallMaxQubitsC128;
    const double state_bytes = (double)(uint64_t(1) << c->n) * amp_bytes(dtype);
    std::vector<HostOp> merged = merge_ops(forward_ops(c->gates));
    if (c->n <= small_cap) {
        if (passes) *passes = 1;
        if (phases) *phases = 1;
        if (ops) *ops = (int)merged.size();
        if (hbm_bytes) *hbm_bytes = 2.0 * state_bytes;
        return DMV_OK;
    }
    Plan plan;
    DMV_TRY(build_plan(merged, tile_params(c->n, dtype, false, default_layout(c->n, dtype), &c->tune_fwd), &plan));
    if (passes) *passes = (int)plan.passes.size();
    if (phases) *phases = (int)plan.phases.size();
    if (ops) *ops = (int)plan.ops.size();
    if (hbm_bytes) *hbm_bytes = 2.0 * state_bytes * (double)plan.passes.size();
    return DMV_OK;
}

int dmv_circuit_plan_ops(dmv_circuit *c, int dtype, int tile_qubits, int *kind, int *target, uint64_t *control_mask,
                         double *m, int *pass, int capacity, int *count)
{
    CHECK_CIRCUIT(c);
    PlanParams pp = tile_params(c->n, dtype, false, default_layout(c->n, dtype), &c->tune_fwd);
    if (tile_qubits > 0) {
        pp.k = tile_qubits;
        pp.tlog = tile_qubits >= 10 ? 8 : 6;
        pp.rlog = tile_qubits - pp.tlog;
    }
    Plan plan;
    DMV_TRY(build_plan(merge_ops(forward_ops(c->gates)), pp, &plan));
    std::vector<int> pass_of;
    std::vector<HostOp> g = plan_global_ops(plan, &pass_of);
    if (count)
        *count = (int)g.size();
    for (int i = 0; i < (int)g.size() && i < capacity; i++) {
        if (kind) kind[i] = (int)g[i].kind;
        if (target) target[i] = g[i].target;
        if (control_mask) control_mask[i] = g[i].ctrl;
        if (m) std::memcpy(m + 8 * i, g[i].m, 8 * sizeof(double));
        if (pass) pass[i] = pass_of[i];
    }
    return DMV_OK;
}

static void json_u64s(std::string &o, const char *key, const uint64_t *a, int n)
{
    o += "\"";
    o += key;
    o += "\":[";
    for (int i = 0; i < n; i++) {
        if (i)
            o += ",";
        o += std::to_string(a[i]);
    }
    o += "]";
}

static void json_u32s(std::string &o, const char *key, const uint32_t *a, int n)
{
    o += "\"";
    o += key;
    o += "\":[";
    for (int i = 0; i < n; i++) {
        if (i)
            o += ",";
        o += std::to_string(a[i]);
    }
    o += "]";
}

static void json_layout(std::string &o, const char *key, const DevLayout &L, const PlanParams &pp)
{
    o += "\"";
    o += key;
    o += "\":{";
    json_u32s(o, "lreg", L.lreg, 1 << pp.rlog);
    o += ",";
    json_u32s(o, "lswz", L.lswz, 1 << pp.rlog);
    o += ",";
    json_u32s(o, "thr", L.thr, pp.tlog);
    o += "}";
}

int dmv_circuit_plan_json(dmv_circuit *c, int dtype, int adjoint, int tile_qubits, int threads_log2,
                          int max_gen_per_pass, char *buffer, uint64_t capacity, uint64_t *length)
{
    CHECK_CIRCUIT(c);
    PlanParams pp = tile_params(c->n, dtype, adjoint != 0, default_layout(c->n, dtype),
                                adjoint ? &c->tune_adj : &c->tune_fwd);
    if (tile_qubits > 0) {
        pp.k = tile_qubits;
        pp.tlog = threads_log2 > 0 ? threads_log2 : (tile_qubits >= 10 ? 8 : 6);
        pp.rlog = pp.k - pp.tlog;
        pp.half_lds = !adjoint && tune_value("half_lds", 1) != 0 && tile_half_exists(dtype, pp.rlog, pp.tlog) ? 1 : 0;
    }
    if (max_gen_per_pass > 0) {
        pp.max_gen_per_pass = max_gen_per_pass;
        pp.gen_auto = 0;
    }
    std::vector<HostOp> ops = merge_ops(adjoint ? adjoint_ops(c->gates) : forward_ops(c->gates));
    Plan plan;
    DMV_TRY(build_plan(ops, pp, &plan));
    std::string o = "{";
    o += "\"n\":" + std::to_string(pp.n) + ",\"k\":" + std::to_string(pp.k) + ",\"rlog\":" + std::to_string(pp.rlog) +
         ",\"tlog\":" + std::to_string(pp.tlog) + ",\"n_slots\":" + std::to_string(plan.n_slots) +
         ",\"layout\":[" + std::to_string(pp.layout.shift) + "," + std::to_string(pp.layout.blk) + "," +
         std::to_string(pp.layout.fold) + "],\"passes\":[";
    for (size_t i = 0; i < plan.passes.size(); i++) {
        const DevPass &P = plan.passes[i];
        if (i)
            o += ",";
        o += "{";
        json_u32s(o, "qpos", P.qpos, pp.k);
        o += ",";
        json_u64s(o, "greg_ld", P.greg_ld, 1 << pp.rlog);
        o += ",";
        json_u64s(o, "greg_st", P.greg_st, 1 << pp.rlog);
        o += ",";
        json_u32s(o, "gthr_ld", P.gthr_ld, pp.tlog);
        o += ",";
        json_u32s(o, "gthr_st", P.gthr_st, pp.tlog);
        o += ",";
        json_layout(o, "lay_ld", P.lay_ld, pp);
        o += ",";
        json_layout(o, "lay_st", P.lay_st, pp);
        o += ",";
        json_u32s(o, "st_src_lswz", P.st_src_lswz, 1 << pp.rlog);
        o += ",";
        json_u32s(o, "st_dst_lswz", P.st_dst_lswz, 1 << pp.rlog);
        o += ",\"transpose_st\":" + std::to_string(P.transpose_st) + ",\"split\":" + std::to_string(P.split) +
             ",\"lds_half\":" + std::to_string(P.lds_half) + ",\"phases\":[";
        for (uint32_t p = P.phase_begin; p < P.phase_end; p++) {
            const DevPhase &D = plan.phases[p];
            if (p != P.phase_begin)
                o += ",";
            o += "{";
            json_layout(o, "lay", D.lay, pp);
            o += ",";
            json_u32s(o, "src_lswz", D.src_lswz, 1 << pp.rlog);
            o += ",";
            json_u32s(o, "dst_lswz", D.dst_lswz, 1 << pp.rlog);
            o += ",\"transpose_in\":" + std::to_string(D.transpose_in) + ",\"ops\":[";
            for (uint32_t q = D.op_begin; q < D.op_end; q++) {
                const DevOp &op = plan.ops[q];
                if (q != D.op_begin)
                    o += ",";
                char buf[512];
                snprintf(buf, sizeof(buf),
                         "{\"kind\":%u,\"target\":%u,\"rmask\":%u,\"tbit\":%u,\"slot\":%d,\"en\":%u,\"fl\":%u,"
                         "\"ctrl_thr\":%u,\"plain\":%u,\"ctrl\":%llu,\"m\":[%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g]}",
                         op.kind, op.target, op.rmask, op.tbit, op.slot, op.en, op.fl, op.ctrl_thr, op.plain,
                         (unsigned long long)op.ctrl, op.m[0], op.m[1], op.m[2], op.m[3], op.m[4], op.m[5], op.m[6],
                         op.m[7]);
                o += buf;
            }
            o += "]}";
        }
        o += "]}";
    }
    o += "]}";
    if (length)
        *length = o.size() + 1;
    if (buffer && capacity > 0) {
        const uint64_t n = std::min<uint64_t>(capacity - 1, o.size());
        std::memcpy(buffer, o.data(), n);
        buffer[n] = '\0';
    }
    return DMV_OK;
}

int dmv_batch_forward_adjoint(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                              const int *gate_control, const int *gate_target, const double *gate_param,
                              const int *coef_offset, const double *coef, const int *param_offset, double *energy,
                              double *grad)
{
    return batch_forward_adjoint(count, dtype, n, gate_offset, gate_kind, gate_control, gate_target, gate_param,
                                 coef_offset, coef, param_offset, energy, grad, nullptr, nullptr);
}

int dmv_batch_autotune(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                       const int *gate_control, const int *gate_target, const double *gate_param,
                       const int *coef_offset, const double *coef, const int *param_offset, int *chosen_tile_qubits)
{
    return autotune_batch(count, dtype, n, gate_offset, gate_kind, gate_control, gate_target, gate_param, coef_offset,
                          coef, param_offset, chosen_tile_qubits);
}

// --------------------------------------------------------------- profiling --
int dmv_profile_enable(int on) { return profile_enable(on); }
int dmv_profile_reset(void) { return profile_reset(); }
int dmv_profile_get(int kc, int64_t *launches, double *total_ms, double *bytes)
{
    return profile_get(kc, launches, total_ms, bytes);
}

int dmv_copy_benchmark(uint64_t bytes, int repeats, double *ms_per_copy)
{
    DMV_CHECK_ARG(bytes >= 16 && repeats >= 1 && ms_per_copy, "bad copy benchmark arguments");
    DMV_TRY(ensure_device());
    void *a = nullptr, *b = nullptr;
    DMV_TRY(alloc(&a, bytes));
    int rc = alloc(&b, bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == DMV_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        set_error("hipEventCreate failed");
        rc = DMV_ERR_HIP;
    }
    if (rc == DMV_OK)
        rc = launch_copy(b, a, bytes);  // warm-up
    if (rc == DMV_OK) {
        (void)hipEventRecord(e0, stream());
        for (int r = 0; r < repeats && rc == DMV_OK; r++) {
            prof_begin(DMV_KC_COPY);
            rc = launch_copy(b, a, bytes);
            prof_end(DMV_KC_COPY, 2.0 * (double)bytes);
        }
        (void)hipEventRecord(e1, stream());
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *ms_per_copy = ms / repeats;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    release(a);
    release(b);
    return rc;
}

int dmv_synchronize(void) { return synchronize(); }

// ------------------------------------------------------------------ tuning --
int dmv_tune_set(const char *name, int64_t value)
{
    DMV_CHECK_ARG(name != nullptr, "null name");
    tune_set(name, value);
    return DMV_OK;
}

int dmv_tune_get(const char *name, int64_t *value)
{
    DMV_CHECK_ARG(name && value, "null argument");
    *value = tune_value(name, -1);
    return DMV_OK;
}

int dmv_autotune(dmv_circuit *c, dmv_state *s, int *chosen_tile_qubits)
{
    CHECK_CIRCUIT(c);
    CHECK_STATE(s);
    return autotune_forward(c, s, chosen_tile_qubits);
}

int dmv_tune_forget(void)
{
    tune_table_reset();
    return DMV_OK;
}

int dmv_autotune_adjoint(dmv_circuit *c, int dtype, int *chosen_tile_qubits)
{
    CHECK_CIRCUIT(c);
    return autotune_adjoint(c, dtype, chosen_tile_qubits);
}

int dmv_circuit_tile_shape(dmv_circuit *c, int dtype, int adjoint, int *tile_qubits, int *threads_log2,
                           int *max_gen_per_pass, int *nontemporal)
{
    CHECK_CIRCUIT(c);
    DMV_CHECK_ARG(dtype == DMV_C128 || dtype == DMV_C64, "dtype must be DMV_C128 or DMV_C64");
    const PlanParams pp = tile_params(c->n, dtype, adjoint != 0, default_layout(c->n, dtype),
                                      adjoint ? &c->tune_adj : &c->tune_fwd);
    if (tile_qubits)
        *tile_qubits = pp.k;
    if (threads_log2)
        *threads_log2 = pp.tlog;
    if (max_gen_per_pass) {
        *max_gen_per_pass = pp.max_gen_per_pass;
        if (pp.gen_auto && c->n > pp.k) {
            // chosen per circuit: the budget its forward plan takes
            Plan plan;
            DMV_TRY(build_plan(merge_ops(forward_ops(c->gates)), pp, &plan));
            *max_gen_per_pass = plan.max_gen;
        }
    }
    if (nontemporal)
        *nontemporal = pp.nontemporal;
    return DMV_OK;
}

}  // extern "C"
