// Circuit lowering and the fused-pass scheduler.  Host C++ only.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dmv {

bool gate_is_parametric(int kind)
{
    return kind == DMV_GATE_RX || kind == DMV_GATE_RY || kind == DMV_GATE_RZ;
}

int gate_num_params(int kind)
{
    if (gate_is_parametric(kind))
        return 1;
    if (kind == DMV_GATE_U3)
        return 3;
    return 0;
}

// Matrices of src/qubit_backend/gates.rs (H :98-119, CNOT :170-174,
// X :224-233, Y :289-298, Z :353-362, RX :418-439, RY :490-511, RZ :562-577,
// S :627-636, T :685-697, Identity :40-49).  U3 and CZ are additions.
int gate_matrix(int kind, const double *p, double m[8])
{
    const double t = p ? p[0] : 0.0;
    for (int i = 0; i < 8; i++)
        m[i] = 0.0;
    switch (kind) {
    case DMV_GATE_ID:
        m[0] = 1.0; m[6] = 1.0;
        return DMV_OK;
    case DMV_GATE_H:
        m[0] = 1.0 / std::sqrt(2.0); m[2] = 1.0 / std::sqrt(2.0);
        m[4] = 1.0 / std::sqrt(2.0); m[6] = -1.0 / std::sqrt(2.0);
        return DMV_OK;
    case DMV_GATE_X:
    case DMV_GATE_CNOT:
        m[2] = 1.0; m[4] = 1.0;
        return DMV_OK;
    case DMV_GATE_Y:
        m[3] = -1.0; m[5] = 1.0;
        return DMV_OK;
    case DMV_GATE_Z:
    case DMV_GATE_CZ:
        m[0] = 1.0; m[6] = -1.0;
        return DMV_OK;
    case DMV_GATE_RX:
        m[0] = std::cos(t / 2.0); m[3] = -std::sin(t / 2.0);
        m[5] = -std::sin(t / 2.0); m[6] = std::cos(t / 2.0);
        return DMV_OK;
    case DMV_GATE_RY:
        m[0] = std::cos(t / 2.0); m[2] = -std::sin(t / 2.0);
        m[4] = std::sin(t / 2.0); m[6] = std::cos(t / 2.0);
        return DMV_OK;
    case DMV_GATE_RZ:
        m[0] = std::cos(-t / 2.0); m[1] = std::sin(-t / 2.0);
        m[6] = std::cos(t / 2.0); m[7] = std::sin(t / 2.0);
        return DMV_OK;
    case DMV_GATE_S:
        m[0] = 1.0; m[7] = 1.0;
        return DMV_OK;
    case DMV_GATE_T:
        m[0] = 1.0; m[6] = std::sqrt(2.0) / 2.0; m[7] = std::sqrt(2.0) / 2.0;
        return DMV_OK;
    case DMV_GATE_U3: {
        const double th = p[0], ph = p[1], la = p[2];
        const double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
        m[0] = c;
        m[2] = -std::cos(la) * s; m[3] = -std::sin(la) * s;
        m[4] = std::cos(ph) * s; m[5] = std::sin(ph) * s;
        m[6] = std::cos(ph + la) * c; m[7] = std::sin(ph + la) * c;
        return DMV_OK;
    }
    default:
        set_error("unknown gate kind %d", kind);
        return DMV_ERR_ARG;
    }
}

uint32_t gate_class(int kind)
{
    switch (kind) {
    case DMV_GATE_X:
    case DMV_GATE_CNOT:
        return OP_SWAP;
    case DMV_GATE_Z:
    case DMV_GATE_CZ:
    case DMV_GATE_RZ:
    case DMV_GATE_S:
    case DMV_GATE_T:
    case DMV_GATE_ID:
        return OP_DIAG;
    default:
        return OP_GEN;
    }
}

static HostOp gate_op(const Gate &g)
{
    HostOp op;
    op.kind = gate_class(g.kind);
    op.target = g.target;
    op.ctrl = g.control >= 0 ? (uint64_t(1) << g.control) : 0;
    gate_matrix(g.kind, g.p, op.m);
    op.slot = -1;
    return op;
}

std::vector<HostOp> forward_ops(const std::vector<Gate> &gates)
{
    std::vector<HostOp> out;
    out.reserve(gates.size());
    for (const Gate &g : gates)
        if (!g.observable)
            out.push_back(gate_op(g));
    return out;
}

// Conjugate transpose of a 2x2 stored as m00 m01 m10 m11 (re, im).
static void dagger(const double m[8], double d[8])
{
    d[0] = m[0]; d[1] = -m[1];   // d00 = conj(m00)
    d[2] = m[4]; d[3] = -m[5];   // d01 = conj(m10)
    d[4] = m[2]; d[5] = -m[3];   // d10 = conj(m01)
    d[6] = m[6]; d[7] = -m[7];   // d11 = conj(m11)
}

static void matmul2(const double a[8], const double b[8], double c[8]);

// Reverse sweep with every IP op moved to the start of its qubit's run of
// uncontrolled one-qubit ops.  Within such a run the state at the IP's own
// place is U psi, U the product of the run's ops before it (both psi and
// lambda carry it), so Im<U lambda|P|U psi> = Im<lambda|U^dag P U|psi> with
// U^dag P U = ax X + ay Y + az Z (Hermitian, traceless, real coefficients).
// Ops on other qubits commute with all of this.  The run's ops then merge
// into one 2x2 (merge_ops) and the run's IP ops share one pass over the
// registers: an RY/RZ layer's reverse step is one GEN and one IP group
// instead of two IP ops, a GEN and a diagonal.
static std::vector<HostOp> hoist_ips(const std::vector<HostOp> &ops, int n)
{
    struct Run {
        bool open = false;
        double U[8];
        int group = -1;
    };
    std::vector<Run> run(n);
    std::vector<std::vector<HostOp>> groups;
    std::vector<int> where;  // per output entry: group index, or -1 for an op
    std::vector<HostOp> body;
    auto open_run = [&](int q) {
        Run &R = run[q];
        if (R.open)
            return;
        R.open = true;
        const double I[8] = {1, 0, 0, 0, 0, 0, 1, 0};
        std::memcpy(R.U, I, sizeof(I));
        R.group = (int)groups.size();
        groups.emplace_back();
        where.push_back(R.group);
        body.emplace_back();
    };
    for (const HostOp &op : ops) {
        const int q = op.target;
        if (op_is_ip(op.kind)) {
            open_run(q);
            Run &R = run[q];
            // A = U^dag P U with P = ax X + ay Y + az Z from op.m[0..2]
            const double P[8] = {op.m[2], 0.0, op.m[0], -op.m[1], op.m[0], op.m[1], -op.m[2], 0.0};
            double Ud[8], t[8], A[8];
            Ud[0] = R.U[0]; Ud[1] = -R.U[1]; Ud[2] = R.U[4]; Ud[3] = -R.U[5];
            Ud[4] = R.U[2]; Ud[5] = -R.U[3]; Ud[6] = R.U[6]; Ud[7] = -R.U[7];
            matmul2(P, R.U, t);
            matmul2(Ud, t, A);
            HostOp ip = op;
            std::memset(ip.m, 0, sizeof(ip.m));
            ip.m[0] = A[4];  // A10 = ax + i ay
            ip.m[1] = A[5];
            ip.m[2] = A[0];  // A00 = az
            groups[R.group].push_back(ip);
            continue;
        }
        if (op.ctrl == 0) {
            open_run(q);
            Run &R = run[q];
            double nu[8];
            matmul2(op.m, R.U, nu);  // op acts after the run so far
            std::memcpy(R.U, nu, sizeof(nu));
        } else {
            run[q].open = false;
            for (int c = 0; c < n; c++)
                if ((op.ctrl >> c) & 1)
                    run[c].open = false;
        }
        where.push_back(-1);
        body.push_back(op);
    }
    std::vector<HostOp> out;
    out.reserve(ops.size());
    for (size_t i = 0; i < where.size(); i++) {
        if (where[i] < 0) {
            out.push_back(body[i]);
            continue;
        }
        std::vector<HostOp> &G = groups[where[i]];
        // one kind per group, so that the kernel computes its sums once
        bool general = false;
        for (const HostOp &ip : G)
            general = general || ip.m[0] != 0.0 || ip.m[1] != 0.0;
        for (HostOp &ip : G) {
            ip.kind = general ? OP_IPG : OP_IPZ;
            out.push_back(ip);
        }
    }
    return out;
}

std::vector<HostOp> adjoint_ops(const std::vector<Gate> &gates)
{
    // parameter index of every gate
    std::vector<int> pidx(gates.size(), -1);
    int np = 0, n = 1;
    for (size_t i = 0; i < gates.size(); i++) {
        if (gate_is_parametric(gates[i].kind))
            pidx[i] = np++;
        n = std::max(n, std::max(gates[i].target, gates[i].control) + 1);
    }
    std::vector<HostOp> out;
    out.reserve(2 * gates.size());
    for (size_t j = gates.size(); j-- > 0;) {
        const Gate &g = gates[j];
        if (g.observable)
            continue;
        if (pidx[j] >= 0) {
            // dE/dtheta = Im<lambda|P|psi> for e^{-i theta P / 2}, P on the target
            HostOp ip;
            ip.kind = g.kind == DMV_GATE_RZ ? OP_IPZ : OP_IPG;
            ip.target = g.target;
            ip.ctrl = 0;
            std::memset(ip.m, 0, sizeof(ip.m));
            ip.m[g.kind == DMV_GATE_RX ? 0 : (g.kind == DMV_GATE_RY ? 1 : 2)] = 1.0;
            ip.slot = pidx[j];
            out.push_back(ip);
        }
        HostOp op = gate_op(g);
        double d[8];
        dagger(op.m, d);
        std::memcpy(op.m, d, sizeof(d));
        out.push_back(op);
    }
    return tune_value("adjoint_hoist", 1) ? hoist_ips(out, n) : out;
}

// c = a * b for 2x2 complex matrices.
static void matmul2(const double a[8], const double b[8], double c[8])
{
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            double re = 0.0, im = 0.0;
            for (int l = 0; l < 2; l++) {
                const double ar = a[2 * (2 * i + l)], ai = a[2 * (2 * i + l) + 1];
                const double br = b[2 * (2 * l + j)], bi = b[2 * (2 * l + j) + 1];
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
            }
            c[2 * (2 * i + j)] = re;
            c[2 * (2 * i + j) + 1] = im;
        }
}

static uint32_t merged_class(uint32_t earlier, uint32_t later)
{
    if (earlier == OP_DIAG && later == OP_DIAG)
        return OP_DIAG;
    if (earlier == OP_SWAP && later == OP_SWAP)
        return OP_DIAG;  // X * X = I exactly
    return OP_GEN;
}

std::vector<HostOp> merge_ops(const std::vector<HostOp> &ops)
{
    std::vector<HostOp> out;
    out.reserve(ops.size());
    int pending[64];
    for (int &p : pending)
        p = -1;
    for (const HostOp &op : ops) {
        const bool one_qubit = op.ctrl == 0 && !op_is_ip(op.kind);
        if (one_qubit) {
            const int pi = pending[op.target];
            if (pi >= 0) {
                HostOp &prev = out[pi];
                double c[8];
                matmul2(op.m, prev.m, c);
                std::memcpy(prev.m, c, sizeof(c));
                prev.kind = merged_class(prev.kind, op.kind);
                continue;
            }
            out.push_back(op);
            pending[op.target] = (int)out.size() - 1;
        } else {
            out.push_back(op);
            pending[op.target] = -1;
            for (int q = 0; q < 64; q++)
                if ((op.ctrl >> q) & 1)
                    pending[q] = -1;
        }
    }
    return out;
}

static uint64_t op_qubits(const HostOp &op) { return (uint64_t(1) << op.target) | op.ctrl; }

// Local index of a global bit mask: bit j of the result is qubit qpos[j].
static uint64_t to_local(uint64_t global, const uint32_t *qpos, int k)
{
    uint64_t l = 0;
    for (int j = 0; j < k; j++)
        if ((global >> qpos[j]) & 1)
            l |= uint64_t(1) << j;
    return l;
}

static uint64_t deposit(uint64_t x, const uint32_t *pos, int nbits)
{
    uint64_t r = 0;
    for (int j = 0; j < nbits; j++)
        if ((x >> j) & 1)
            r |= uint64_t(1) << pos[j];
    return r;
}

// Builds the register/thread layout of a phase whose register bits are `rs`.
// coalesced: thread bits 0..low_bits-1 are local bits 0..low_bits-1 (they
// must not be register bits); otherwise thread bits 0..swz_bits-1 get local
// bits with distinct residues mod swz_bits, which keeps 128-bit LDS accesses
// of consecutive lanes in distinct banks under the swizzle of the kernels.
// top >= 0: local bit `top` (not in rs) takes the highest thread bit.
static DevLayout make_layout(uint32_t rs, const PlanParams &pp, bool coalesced, int top)
{
    DevLayout L;
    std::memset(&L, 0, sizeof(L));
    uint32_t regbits[kMaxRegs];
    int nr = 0;
    for (int b = 0; b < pp.k; b++)
        if ((rs >> b) & 1)
            regbits[nr++] = (uint32_t)b;
    for (int r = 0; r < (1 << pp.rlog); r++)
        L.lreg[r] = (uint32_t)deposit((uint64_t)r, regbits, nr);
    std::vector<int> tb;
    for (int b = 0; b < pp.k; b++)
        if (!((rs >> b) & 1) && b != top)
            tb.push_back(b);
    std::vector<int> order;
    if (coalesced) {
        for (int b = 0; b < pp.low_bits; b++)
            order.push_back(b);
    } else {
        for (int c = 0; c < pp.swz_bits && (int)order.size() < (int)tb.size(); c++)
            for (int b : tb)
                if (b % pp.swz_bits == c && std::find(order.begin(), order.end(), b) == order.end()) {
                    order.push_back(b);
                    break;
                }
    }
    for (int b : tb)
        if (std::find(order.begin(), order.end(), b) == order.end())
            order.push_back(b);
    if (top >= 0)
        order.push_back(top);
    for (int j = 0; j < pp.tlog; j++)
        L.thr[j] = (uint32_t)order[j];
    const uint32_t elem = pp.swz_bits == 3 ? 16u : 8u;
    for (int r = 0; r < (1 << pp.rlog); r++)
        L.lswz[r] = (pp.swz_bits == 3 ? swz_index<3>(L.lreg[r]) : swz_index<4>(L.lreg[r])) * elem;
    return L;
}

// Register bit of a local bit in a layout, or kNoReg.
static uint32_t register_bit(const DevLayout &L, const PlanParams &pp, uint32_t lbit)
{
    for (int j = 0; j < pp.rlog; j++)
        if (L.lreg[1 << j] == (1u << lbit))
            return (uint32_t)j;
    return kNoReg;
}

// Thread bit of a local bit in a layout, or kNoReg.
static uint32_t thread_bit(const DevLayout &L, const PlanParams &pp, uint32_t lbit)
{
    for (int j = 0; j < pp.tlog; j++)
        if (L.thr[j] == lbit)
            return (uint32_t)j;
    return kNoReg;
}

static bool layouts_equal(const DevLayout &a, const DevLayout &b, const PlanParams &pp)
{
    for (int r = 0; r < (1 << pp.rlog); r++)
        if (a.lreg[r] != b.lreg[r])
            return false;
    for (int j = 0; j < pp.tlog; j++)
        if (a.thr[j] != b.thr[j])
            return false;
    return true;
}

// Pads a register set to rlog bits, preferring the highest local bits that
// are not kept in lanes for coalescing, and never taking the bits of `avoid`.
static uint32_t pad_regs(uint32_t rs, const PlanParams &pp, uint32_t avoid)
{
    for (int b = pp.k - 1; b >= pp.low_bits && popcount64(rs) < pp.rlog; b--)
        if (!((avoid >> b) & 1))
            rs |= 1u << b;
    for (int b = 0; b < pp.k && popcount64(rs) < pp.rlog; b++)
        if (!((avoid >> b) & 1))
            rs |= 1u << b;
    return rs;
}

// Registers p with the highest set bit of partner mask m clear: the first
// member of each pair (p, p ^ m) as the kernels enumerate pairs.
static uint32_t first_members(uint32_t m, int nr)
{
    uint32_t hb = 1;
    while ((hb << 1) <= m)
        hb <<= 1;
    uint32_t f = 0;
    for (int p = 0; p < nr; p++)
        if (!(p & hb))
            f |= 1u << p;
    return f;
}

static int make_pass(uint64_t Q, const std::vector<HostOp> &taken, const PlanParams &pp, Plan *plan)
{
    DevPass P;
    std::memset(&P, 0, sizeof(P));
    P.k = (uint32_t)pp.k;
    int nq = 0;
    for (int q = 0; q < pp.n; q++)
        if ((Q >> q) & 1)
            P.qpos[nq++] = (uint32_t)q;
    if (nq != pp.k) {
        set_error("internal: pass has %d qubits, expected %d", nq, pp.k);
        return DMV_ERR_STATE;
    }
    std::vector<HostOp> loc(taken.size());
    for (size_t i = 0; i < taken.size(); i++) {
        loc[i] = taken[i];
        const uint64_t lt = to_local(uint64_t(1) << taken[i].target, P.qpos, pp.k);
        loc[i].target = __builtin_ctzll(lt);
        loc[i].ctrl = to_local(taken[i].ctrl, P.qpos, pp.k);
    }

    // Register phases: greedy, keeping op order wherever ops share a qubit.
    struct Ph {
        uint32_t rs;
        std::vector<int> ops;
    };
    std::vector<Ph> phs;
    std::vector<char> used(loc.size(), 0);
    size_t left = loc.size();
    while (left > 0) {
        Ph ph{0u, {}};
        uint64_t blocked = 0;
        for (size_t i = 0; i < loc.size(); i++) {
            if (used[i])
                continue;
            const uint64_t qs = (uint64_t(1) << loc[i].target) | loc[i].ctrl;
            if (qs & blocked) {
                blocked |= qs;
                continue;
            }
            if (op_needs_register(loc[i].kind)) {
                const uint32_t tb = 1u << loc[i].target;
                if (!(ph.rs & tb)) {
                    if (popcount64(ph.rs) >= pp.rlog) {
                        blocked |= qs;
                        continue;
                    }
                    ph.rs |= tb;
                }
            }
            ph.ops.push_back((int)i);
            used[i] = 1;
            left--;
        }
        if (ph.ops.empty()) {
            set_error("internal: scheduler made no progress");
            return DMV_ERR_STATE;
        }
        phs.push_back(std::move(ph));
    }
    const uint32_t lowmask = (1u << pp.low_bits) - 1;
    // Split transposes: the highest local bit above the lane-coalesced and
    // swizzle-group bits that no op of the pass needs in a register stays a
    // thread bit, at the top, in every layout of the pass (wave-uniform halves).
    int top = -1;
    if (pp.half_lds && pp.tlog >= 7) {
        uint32_t targets = 0;
        for (const Ph &ph : phs)
            targets |= ph.rs;
        for (int b = pp.k - 1; b >= std::max(pp.low_bits, pp.swz_bits) && top < 0; b--)
            if (!((targets >> b) & 1))
                top = b;
    }
    const uint32_t avoid = top >= 0 ? 1u << top : 0u;
    for (Ph &ph : phs)
        ph.rs = pad_regs(ph.rs, pp, avoid);

    // Layouts: load/store layouts keep the low qubits in lanes.  By default
    // the first/last phase's layout is reused when it keeps the low qubits in
    // lanes.  With memory-ordered layouts (tunable) they are always the
    // canonical layout: lanes and waves take the tile's lowest qubits in
    // ascending order and registers its highest, and a phase whose register
    // set differs pays an LDS transpose instead (measured slower overall on
    // the gapped state layout).
    const uint32_t default_rs = pad_regs(0u, pp, avoid);
    const bool mem_order = tune_value("memory_layouts", 0) != 0;
    const bool first_coalesced = (phs.front().rs & lowmask) == 0;
    const bool last_coalesced = (phs.back().rs & lowmask) == 0;
    P.lay_ld = make_layout(first_coalesced && !mem_order ? phs.front().rs : default_rs, pp, true, top);
    const uint32_t elem_log2 = pp.swz_bits == 3 ? 4u : 3u;
    P.split = top >= 0 ? (uint32_t)top + elem_log2 : 0u;
    auto lds_off = [&](uint32_t y) { return P.split ? drop_bit(y, P.split) : y; };
    P.phase_begin = (uint32_t)plan->phases.size();
    const int pass_index = (int)plan->passes.size();
    const int NR = 1 << pp.rlog;
    const uint32_t all_regs = NR >= 32 ? 0xFFFFFFFFu : ((1u << NR) - 1);

    // Register renaming: physical register p holds logical register tab[p]
    // (an affine map over GF(2)).  X gates and CNOTs between register bits
    // only update the map; every later op addresses physical registers
    // through it, and the next transpose or store writes register p to the
    // place of logical register tab[p].
    uint32_t tab[kMaxRegs], inv[kMaxRegs];
    auto reset_tab = [&]() {
        for (int q = 0; q < NR; q++)
            tab[q] = (uint32_t)q;
    };
    auto invert = [&]() {
        for (int q = 0; q < NR; q++)
            inv[tab[q]] = (uint32_t)q;
    };
    reset_tab();
    DevLayout prev = P.lay_ld;
    for (size_t p = 0; p < phs.size(); p++) {
        DevPhase D;
        std::memset(&D, 0, sizeof(D));
        const bool coal = (p == 0 && first_coalesced) ||
                          (p + 1 == phs.size() && last_coalesced && !(p == 0 && !first_coalesced));
        D.lay = make_layout(phs[p].rs, pp, coal && (phs[p].rs & lowmask) == 0, top);
        D.transpose_in = layouts_equal(D.lay, prev, pp) ? 0u : 1u;
        if (D.transpose_in) {
            for (int q = 0; q < NR; q++) {
                D.src_lswz[q] = lds_off(prev.lswz[tab[q]]);
                D.dst_lswz[q] = lds_off(D.lay.lswz[q]);
            }
            reset_tab();
        }
        D.op_begin = (uint32_t)plan->ops.size();
        for (int i : phs[p].ops) {
            const HostOp &h = loc[i];
            {
                HostOp g = h;
                g.target = (int)P.qpos[h.target];
                g.ctrl = deposit(h.ctrl, P.qpos, pp.k);
                plan->logical.push_back(g);
                plan->logical_pass.push_back(pass_index);
            }
            const uint32_t J = register_bit(D.lay, pp, (uint32_t)h.target);
            uint32_t creg = 0, cthr = 0;
            for (int b = 0; b < pp.k; b++) {
                if (!((h.ctrl >> b) & 1))
                    continue;
                const uint32_t rb = register_bit(D.lay, pp, (uint32_t)b);
                if (rb != kNoReg)
                    creg |= 1u << rb;
                else
                    cthr |= 1u << thread_bit(D.lay, pp, (uint32_t)b);
            }
            if (op_needs_register(h.kind) && J == kNoReg) {
                set_error("internal: op target not a register bit");
                return DMV_ERR_STATE;
            }
            // relabel-only X / CNOT: the map stays affine with at most one register control
            if (h.kind == OP_SWAP && cthr == 0 && popcount64(creg) <= 1 && tune_value("register_renaming", 1)) {
                for (int q = 0; q < NR; q++)
                    if ((tab[q] & creg) == creg)
                        tab[q] ^= 1u << J;
                continue;
            }
            invert();
            DevOp op;
            std::memset(&op, 0, sizeof(op));
            op.kind = h.kind;
            op.target = (uint32_t)h.target;
            op.ctrl = h.ctrl;
            op.slot = h.slot;
            op.ctrl_thr = cthr;
            std::memcpy(op.m, h.m, sizeof(op.m));
            op.tbit = J == kNoReg ? thread_bit(D.lay, pp, (uint32_t)h.target) : kNoReg;
            op.en = 0;
            for (int q = 0; q < NR; q++)
                if ((tab[q] & creg) == creg)
                    op.en |= 1u << q;
            if (J != kNoReg) {
                op.fl = 0;
                for (int q = 0; q < NR; q++)
                    if ((tab[q] >> J) & 1)
                        op.fl |= 1u << q;
                // partner of physical q: the physical register of logical tab[q] ^ e_J
                op.rmask = inv[tab[0] ^ (1u << J)] ^ inv[tab[0]];
                if (!op_needs_register(h.kind))
                    op.rmask = op.fl;  // register-target diagonal: registers whose target bit is 1
            } else {
                op.rmask = 0;
                op.fl = 0;
            }
            if (h.kind == OP_GEN && J != kNoReg)
                op.plain = (op.en == all_regs && (op.fl & first_members(op.rmask, NR)) == 0) ? 1u : 0u;
            if (h.kind == OP_DIAG)
                op.plain = diag_is_sign_flip(op.m) ? 1u : 0u;
            if (op_is_ip(h.kind) && h.slot + 1 > plan->n_slots)
                plan->n_slots = h.slot + 1;
            if (op_is_ip(h.kind) && plan->ops.size() > D.op_begin) {
                // an IP op right after one of the same kind and masks reuses its sums
                const DevOp &prev = plan->ops.back();
                op.plain = prev.kind == op.kind && prev.target == op.target && prev.tbit == op.tbit &&
                                   prev.rmask == op.rmask && prev.fl == op.fl
                               ? 1u
                               : 0u;
            }
            plan->ops.push_back(op);
            plan->op_pass.push_back(pass_index);
        }
        D.op_end = (uint32_t)plan->ops.size();
        plan->phases.push_back(D);
        prev = D.lay;
    }
    P.phase_end = (uint32_t)plan->phases.size();
    const DevLayout last = prev;
    const bool store_direct = mem_order ? layouts_equal(last, make_layout(default_rs, pp, true, top), pp)
                                        : last_coalesced &&
                                              layouts_equal(last, make_layout(phs.back().rs, pp, true, top), pp);
    P.lay_st = store_direct ? last : make_layout(default_rs, pp, true, top);
    P.transpose_st = layouts_equal(P.lay_st, last, pp) ? 0u : 1u;
    if (P.transpose_st) {
        for (int q = 0; q < NR; q++) {
            P.st_src_lswz[q] = lds_off(last.lswz[tab[q]]);
            P.st_dst_lswz[q] = lds_off(P.lay_st.lswz[q]);
        }
        reset_tab();
    }
    // a pass without transposes needs no LDS at all
    bool any_transpose = P.transpose_st != 0;
    for (uint32_t p = P.phase_begin; p < (uint32_t)plan->phases.size(); p++)
        any_transpose = any_transpose || plan->phases[p].transpose_in;
    P.lds_half = (P.split || !any_transpose) && pp.half_lds ? 1u : 0u;
    // global offsets of the load/store layouts (the store through the final map)
    for (int r = 0; r < NR; r++) {
        // element offsets: the kernels add them to the element of tile base | thread offset
        P.greg_ld[r] = elem(deposit(P.lay_ld.lreg[r], P.qpos, pp.k), pp.layout);
        P.greg_st[r] = elem(deposit(P.lay_st.lreg[tab[r]], P.qpos, pp.k), pp.layout);
    }
    for (int j = 0; j < pp.tlog; j++) {
        P.gthr_ld[j] = P.qpos[P.lay_ld.thr[j]];
        P.gthr_st[j] = P.qpos[P.lay_st.thr[j]];
    }
    plan->passes.push_back(P);
    return DMV_OK;
}

// Plan cost in units of a plain pass: every pass moves the state once, and
// every phase beyond a pass's first adds compute between barriers (measured
// on MI355X: the 30-qubit layered circuit at budget 12 adds one phase to
// every pass for 1.29x the time; the 33-qubit random circuit saves 7 of 58
// passes for 0.93x).
static double plan_cost(const Plan &plan)
{
    return (double)plan.passes.size() + 0.6 * (double)(plan.phases.size() - plan.passes.size());
}

int build_plan(const std::vector<HostOp> &ops, const PlanParams &pp, Plan *out)
{
    if (pp.gen_auto) {
        // the GEN budget per pass depends on the circuit: plan with 8 and 12
        // and keep the cheaper plan (the smaller budget on a tie)
        PlanParams a = pp, b = pp;
        a.gen_auto = b.gen_auto = 0;
        a.max_gen_per_pass = 8;
        b.max_gen_per_pass = 12;
        Plan pa, pb;
        DMV_TRY(build_plan(ops, a, &pa));
        DMV_TRY(build_plan(ops, b, &pb));
        *out = plan_cost(pb) < plan_cost(pa) ? std::move(pb) : std::move(pa);
        return DMV_OK;
    }
    DMV_CHECK_ARG(pp.k == pp.rlog + pp.tlog, "tile qubits %d != %d + %d", pp.k, pp.rlog, pp.tlog);
    DMV_CHECK_ARG(pp.n >= pp.k && pp.k <= kMaxTileQubits && pp.rlog >= 1 &&
                      (1 << pp.rlog) <= kMaxRegs && pp.tlog <= kMaxThreadBits &&
                      pp.low_bits <= pp.tlog && pp.low_bits + pp.rlog <= pp.k,
                  "invalid tile shape n=%d k=%d rlog=%d tlog=%d", pp.n, pp.k, pp.rlog, pp.tlog);
    Plan plan;
    const uint64_t all = (pp.n >= 64) ? ~uint64_t(0) : ((uint64_t(1) << pp.n) - 1);
    const uint64_t lowmask = (uint64_t(1) << pp.low_bits) - 1;
    for (const HostOp &op : ops) {
        DMV_CHECK_ARG(op.target >= 0 && op.target < pp.n, "target %d out of range", op.target);
        DMV_CHECK_ARG((op.ctrl & ~all) == 0 && !((op.ctrl >> op.target) & 1), "bad control mask");
    }
    // With split transposes, a pass keeps one of its qubits above the lane
    // and swizzle-group bits free of register targets (the split bit): an op
    // that would make every such qubit a register target waits for the next
    // pass.  The lowest qubits are always in the tile, so those candidates
    // are the tile's qubits outside the lowest max(low_bits, swz_bits).
    const int split_keep = pp.half_lds && pp.tlog >= 7 ? pp.k - std::max(pp.low_bits, pp.swz_bits) - 1 : 64;
    const uint64_t split_low = (uint64_t(1) << std::max(pp.low_bits, pp.swz_bits)) - 1;
    std::vector<char> done(ops.size(), 0);
    size_t first = 0;
    while (first < ops.size()) {
        uint64_t Q = lowmask, blocked = 0, treg = 0;
        int gen = 0;
        std::vector<HostOp> taken;
        for (size_t i = first; i < ops.size(); i++) {
            if (done[i])
                continue;
            const uint64_t qs = op_qubits(ops[i]);
            if (qs & blocked) {
                blocked |= qs;
                if ((blocked & all) == all)
                    break;
                continue;
            }
            const uint64_t nq = Q | qs;
            const bool budget = ops[i].kind != OP_GEN || gen < pp.max_gen_per_pass;
            const uint64_t nt = op_needs_register(ops[i].kind) ? (treg | (uint64_t(1) << ops[i].target)) : treg;
            const bool split_ok = popcount64(nt & ~split_low) <= split_keep;
            if (popcount64(nq) <= pp.k && budget && split_ok) {
                Q = nq;
                treg = nt;
                taken.push_back(ops[i]);
                done[i] = 1;
                if (ops[i].kind == OP_GEN)
                    gen++;
            } else {
                blocked |= qs;
                if ((blocked & all) == all)
                    break;
            }
        }
        while (first < ops.size() && done[first])
            first++;
        // fill the tile up to k qubits with the lowest unused qubits
        for (int q = 0; q < pp.n && popcount64(Q) < pp.k; q++)
            Q |= uint64_t(1) << q;
        DMV_TRY(make_pass(Q, taken, pp, &plan));
    }
    // number the IP ops of every pass: pass p's gradient partials are rows
    // ip_begin[p] .. ip_begin[p+1] of the per-pass partial buffer
    plan.ip_begin.assign(plan.passes.size() + 1, 0);
    plan.ip_slots.clear();
    for (size_t p = 0; p < plan.passes.size(); p++) {
        plan.ip_begin[p] = (uint32_t)plan.ip_slots.size();
        for (uint32_t ph = plan.passes[p].phase_begin; ph < plan.passes[p].phase_end; ph++)
            for (uint32_t o = plan.phases[ph].op_begin; o < plan.phases[ph].op_end; o++)
                if (op_is_ip(plan.ops[o].kind)) {
                    plan.ops[o].ip_pos = (uint32_t)plan.ip_slots.size() - plan.ip_begin[p];
                    plan.ip_slots.push_back(plan.ops[o].slot);
                }
    }
    plan.ip_begin[plan.passes.size()] = (uint32_t)plan.ip_slots.size();
    DMV_TRY(validate_plan(plan, pp));
    plan.max_gen = pp.max_gen_per_pass;
    *out = std::move(plan);
    return DMV_OK;
}

static int check_layout(const DevLayout &L, const PlanParams &pp, const char *what)
{
    uint32_t seen = 0;
    for (int j = 0; j < pp.tlog; j++) {
        if (L.thr[j] >= (uint32_t)pp.k || ((seen >> L.thr[j]) & 1)) {
            set_error("internal: %s thread bit %d invalid", what, j);
            return DMV_ERR_STATE;
        }
        seen |= 1u << L.thr[j];
    }
    uint32_t regmask = 0;
    for (int j = 0; j < pp.rlog; j++)
        regmask |= L.lreg[1 << j];
    if ((regmask & seen) != 0 || popcount64(regmask) != pp.rlog || ((regmask | seen) + 1) != (1u << pp.k)) {
        set_error("internal: %s register bits invalid", what);
        return DMV_ERR_STATE;
    }
    const uint32_t elem = pp.swz_bits == 3 ? 16u : 8u;
    for (int r = 0; r < (1 << pp.rlog); r++) {
        uint32_t expect = 0;
        for (int j = 0; j < pp.rlog; j++)
            if ((r >> j) & 1)
                expect |= L.lreg[1 << j];
        const uint32_t sw = (pp.swz_bits == 3 ? swz_index<3>(L.lreg[r]) : swz_index<4>(L.lreg[r])) * elem;
        if (L.lreg[r] != expect || L.lswz[r] != sw || L.lswz[r] >= (elem << pp.k)) {
            set_error("internal: %s register offsets inconsistent", what);
            return DMV_ERR_STATE;
        }
    }
    return DMV_OK;
}

// A transpose's write offsets must be a permutation of its source layout's
// and its read offsets its destination layout's, with the split bit removed
// in a split pass.  There, the split local bit must be the top thread bit of
// both layouts (so each half of the threads exchanges amplitudes only within
// itself) and the offsets must fit half a tile.
static int check_offsets(const uint32_t *src, const DevLayout &A, const uint32_t *dst, const DevLayout &B,
                         uint32_t split, const PlanParams &pp, const char *what)
{
    const int NR = 1 << pp.rlog;
    const uint32_t elem_log2 = pp.swz_bits == 3 ? 4u : 3u;
    const uint32_t limit = (1u << (pp.k + (int)elem_log2)) >> (split ? 1 : 0);
    std::vector<uint32_t> a(src, src + NR), b(NR), c(dst, dst + NR), d(NR);
    for (int r = 0; r < NR; r++) {
        b[r] = split ? drop_bit(A.lswz[r], split) : A.lswz[r];
        d[r] = split ? drop_bit(B.lswz[r], split) : B.lswz[r];
        if (d[r] != c[r] || a[r] >= limit) {
            set_error("internal: %s read offsets inconsistent", what);
            return DMV_ERR_STATE;
        }
    }
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    if (a != b) {
        set_error("internal: %s offsets are not a permutation of the layout", what);
        return DMV_ERR_STATE;
    }
    if (split) {
        const uint32_t w = split - elem_log2;
        if (split < elem_log2 || (int)w < pp.swz_bits || (int)w >= pp.k || pp.tlog < 7 ||
            A.thr[pp.tlog - 1] != w || B.thr[pp.tlog - 1] != w) {
            set_error("internal: %s split bit is not the top thread bit of both layouts", what);
            return DMV_ERR_STATE;
        }
    }
    return DMV_OK;
}

int validate_plan(const Plan &plan, const PlanParams &pp)
{
    const uint64_t n_amp = uint64_t(1) << pp.n;
    for (const DevPass &P : plan.passes) {
        if ((int)P.k != pp.k) {
            set_error("internal: pass k mismatch");
            return DMV_ERR_STATE;
        }
        for (int j = 0; j < pp.k; j++)
            if (P.qpos[j] >= (uint32_t)pp.n || (j > 0 && P.qpos[j] <= P.qpos[j - 1])) {
                set_error("internal: pass qubit list invalid");
                return DMV_ERR_STATE;
            }
        DMV_TRY(check_layout(P.lay_ld, pp, "load layout"));
        DMV_TRY(check_layout(P.lay_st, pp, "store layout"));
        for (int j = 0; j < pp.low_bits; j++)
            if (P.lay_ld.thr[j] != (uint32_t)j || P.lay_st.thr[j] != (uint32_t)j) {
                set_error("internal: load/store layout not coalesced");
                return DMV_ERR_STATE;
            }
        for (int r = 0; r < (1 << pp.rlog); r++)
            if (P.greg_ld[r] >= n_amp || P.greg_st[r] >= n_amp) {
                set_error("internal: register offset out of range");
                return DMV_ERR_STATE;
            }
        for (int j = 0; j < pp.tlog; j++)
            if (P.gthr_ld[j] >= (uint32_t)pp.n || P.gthr_st[j] >= (uint32_t)pp.n) {
                set_error("internal: thread bit position out of range");
                return DMV_ERR_STATE;
            }
        if (P.phase_begin >= P.phase_end || P.phase_end > plan.phases.size()) {
            set_error("internal: phase range invalid");
            return DMV_ERR_STATE;
        }
        for (uint32_t p = P.phase_begin; p < P.phase_end; p++) {
            const DevPhase &D = plan.phases[p];
            DMV_TRY(check_layout(D.lay, pp, "phase layout"));
            if (D.op_begin > D.op_end || D.op_end > plan.ops.size()) {
                set_error("internal: op range invalid");
                return DMV_ERR_STATE;
            }
            if (D.transpose_in)
                DMV_TRY(check_offsets(D.src_lswz, p == P.phase_begin ? P.lay_ld : plan.phases[p - 1].lay,
                                      D.dst_lswz, D.lay, P.split, pp, "phase transpose"));
            for (uint32_t o = D.op_begin; o < D.op_end; o++) {
                const DevOp &op = plan.ops[o];
                const int NR = 1 << pp.rlog;
                const uint32_t regs = NR >= 32 ? 0xFFFFFFFFu : ((1u << NR) - 1);
                if (op.target >= (uint32_t)pp.k || (op.ctrl >> pp.k) != 0 || ((op.ctrl >> op.target) & 1)) {
                    set_error("internal: op bits out of range");
                    return DMV_ERR_STATE;
                }
                if (op.kind != OP_GEN && op.kind != OP_DIAG && op.kind != OP_SWAP && !op_is_ip(op.kind)) {
                    set_error("internal: op kind invalid");
                    return DMV_ERR_STATE;
                }
                if ((op.en & ~regs) || (op.fl & ~regs) || (op.ctrl_thr >> pp.tlog) != 0) {
                    set_error("internal: op masks out of range");
                    return DMV_ERR_STATE;
                }
                // GENs come from uncontrolled gates only; the reverse sweep's
                // kernels apply them to every pair of every lane
                if (op.kind == OP_GEN && (op.ctrl != 0 || op.ctrl_thr != 0 || op.en != regs)) {
                    set_error("internal: controlled GEN op");
                    return DMV_ERR_STATE;
                }
                if (op_is_ip(op.kind)) {
                    const DevOp *prev = o > D.op_begin ? &plan.ops[o - 1] : nullptr;
                    if (op.plain && !(prev && prev->kind == op.kind && prev->target == op.target &&
                                      prev->tbit == op.tbit && prev->rmask == op.rmask && prev->fl == op.fl)) {
                        set_error("internal: IP op reuses sums it has no predecessor for");
                        return DMV_ERR_STATE;
                    }
                } else if (op.plain && op.kind == OP_DIAG) {
                    if (!diag_is_sign_flip(op.m)) {
                        set_error("internal: plain flag on a diagonal other than diag(1, -1)");
                        return DMV_ERR_STATE;
                    }
                } else if (op.plain &&
                           (op.kind != OP_GEN || op.en != regs || (op.fl & first_members(op.rmask, NR)) != 0)) {
                    set_error("internal: plain flag on an op that needs selects");
                    return DMV_ERR_STATE;
                }
                const bool target_reg = register_bit(D.lay, pp, op.target) != kNoReg;
                if (op_needs_register(op.kind)) {
                    if (!target_reg || op.rmask == 0 || op.rmask >= (uint32_t)NR || op.tbit != kNoReg) {
                        set_error("internal: pair op without a register partner");
                        return DMV_ERR_STATE;
                    }
                } else if (target_reg) {
                    if ((op.rmask & ~regs) || op.tbit != kNoReg) {
                        set_error("internal: register diagonal masks out of range");
                        return DMV_ERR_STATE;
                    }
                } else if (op.tbit >= (uint32_t)pp.tlog || D.lay.thr[op.tbit] != op.target) {
                    set_error("internal: thread-bit target inconsistent");
                    return DMV_ERR_STATE;
                }
                for (int j = 0; j < pp.tlog; j++)
                    if (((op.ctrl_thr >> j) & 1) && !((op.ctrl >> D.lay.thr[j]) & 1)) {
                        set_error("internal: thread control not in the control mask");
                        return DMV_ERR_STATE;
                    }
            }
        }
        if (P.transpose_st)
            DMV_TRY(check_offsets(P.st_src_lswz, plan.phases[P.phase_end - 1].lay, P.st_dst_lswz, P.lay_st, P.split,
                                  pp, "store transpose"));
        bool any_transpose = P.transpose_st != 0;
        for (uint32_t p = P.phase_begin; p < P.phase_end; p++)
            any_transpose = any_transpose || plan.phases[p].transpose_in;
        if ((P.lds_half || P.split) && !pp.half_lds) {
            set_error("internal: split transposes in a plan for full-LDS kernels");
            return DMV_ERR_STATE;
        }
        if (P.lds_half && any_transpose && !P.split) {
            set_error("internal: half-LDS pass with a full transpose");
            return DMV_ERR_STATE;
        }
    }
    return DMV_OK;
}

std::vector<HostOp> plan_global_ops(const Plan &plan, std::vector<int> *pass_of_op)
{
    if (pass_of_op)
        *pass_of_op = plan.logical_pass;
    return plan.logical;
}

std::vector<DevOp> global_dev_ops(const std::vector<HostOp> &ops)
{
    std::vector<DevOp> out(ops.size());
    for (size_t i = 0; i < ops.size(); i++) {
        std::memset(&out[i], 0, sizeof(DevOp));
        out[i].kind = ops[i].kind;
        out[i].target = (uint32_t)ops[i].target;
        out[i].ctrl = ops[i].ctrl;
        out[i].slot = ops[i].slot;
        std::memcpy(out[i].m, ops[i].m, sizeof(out[i].m));
    }
    return out;
}

}  // namespace dmv
