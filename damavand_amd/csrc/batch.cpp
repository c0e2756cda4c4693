// Batched forward + adjoint of many small circuits in one persistent launch.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace dmv {

// The batched kernel's tile shape (tile qubits k, log2 threads) and
// workgroups per CU: from the caller's choice `pin`, else the tunables
// batch_tile_qubits / batch_threads_log2 / batch_wgs_per_cu, else the tuning
// table for (qubits of the largest circuit, dtype), else k = 11 on 256
// threads, two workgroups per CU.
static void batch_shape(int nmax, int dtype, const TuneChoice *pin, int *k, int *tlog, int *wgs)
{
    // three workgroups per CU (the kernel keeps to 168 VGPRs for it);
    // measured on the 4096-circuit 12-18 qubit batch: complex128 219 -> 183-193
    // ms and complex64 200 -> 159 ms against two
    *k = 11;
    *tlog = 8;
    *wgs = 3;
    TuneChoice t;
    if (tune_lookup(tune_key(nmax, dtype, "batch"), &t) && t.set()) {
        *k = t.k;
        *tlog = t.tlog > 0 ? t.tlog : *tlog;
        *wgs = t.wgs > 0 ? t.wgs : *wgs;
    }
    *k = (int)tune_value("batch_tile_qubits", *k);
    *tlog = (int)tune_value("batch_threads_log2", *tlog);
    *wgs = (int)tune_value("batch_wgs_per_cu", *wgs);
    if (pin && pin->set()) {
        *k = pin->k;
        *tlog = pin->tlog > 0 ? pin->tlog : *tlog;
        *wgs = pin->wgs > 0 ? pin->wgs : *wgs;
    }
}

int batch_forward_adjoint(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                          const int *gate_control, const int *gate_target, const double *gate_param,
                          const int *coef_offset, const double *coef, const int *param_offset, double *energy,
                          double *grad, const TuneChoice *pin, double *kernel_ms)
{
    DMV_CHECK_ARG(count >= 1, "empty batch");
    DMV_CHECK_ARG(dtype == DMV_C128 || dtype == DMV_C64, "bad dtype");
    DMV_CHECK_ARG(n && gate_offset && gate_kind && gate_control && gate_target && gate_param && coef_offset && coef &&
                      param_offset && energy && grad,
                  "null batch array");
    // the offsets index host arrays here and device buffers in the kernel:
    // gate ranges must be ordered, coefficient ranges non-negative, and the
    // gradient offsets the exclusive prefix sum of the parameter counts
    // (checked once the counts are known, below)
    DMV_CHECK_ARG(gate_offset[0] >= 0, "gate_offset[0] is negative");
    for (int i = 0; i < count; i++) {
        DMV_CHECK_ARG(gate_offset[i + 1] >= gate_offset[i], "gate_offset decreases at circuit %d", i);
        DMV_CHECK_ARG(coef_offset[i] >= 0, "coef_offset[%d] is negative", i);
    }
    int nmax_all = 0;
    for (int i = 0; i < count; i++)
        nmax_all = std::max(nmax_all, n[i]);
    int K, tlog, per_cu;
    batch_shape(nmax_all, dtype, pin, &K, &tlog, &per_cu);
    const int rlog = K - tlog;
    DMV_CHECK_ARG(batch_kernel_exists(dtype, rlog, tlog), "no batched kernel for %d tile qubits on %d threads", K,
                  1 << tlog);
    DMV_CHECK_ARG(per_cu >= 1 && per_cu <= 16, "batch_wgs_per_cu must be in [1, 16], got %d", per_cu);
    const int small_max = K - 1;

    // Per-circuit plans are independent: build them on host threads, then
    // concatenate in circuit order (the result does not depend on threading).
    struct Built {
        int rc = DMV_OK;
        std::string err;
        int np = 0;
        double cost = 0.0;
        bool small = false;
        std::vector<DevOp> fo, ao;
        Plan pf, pa;
    };
    std::vector<Built> built(count);
    PlanParams pp0;
    pp0.k = K;
    pp0.rlog = rlog;
    pp0.tlog = tlog;
    pp0.low_bits = coalesce_bits(dtype);
    pp0.swz_bits = dtype == DMV_C64 ? 4 : 3;
    pp0.max_gen_per_pass = (int)tune_value("max_gen_per_pass", 8);
    auto build_one = [&](int i, Built &B) -> int {
        const int ni = n[i];
        DMV_CHECK_ARG(ni >= 2 && ni <= 24, "batch circuit %d: qubits must be in [2, 24], got %d", i, ni);
        std::vector<Gate> gates;
        for (int g = gate_offset[i]; g < gate_offset[i + 1]; g++) {
            Gate G;
            G.kind = gate_kind[g];
            G.control = gate_control[g];
            G.target = gate_target[g];
            G.p[0] = gate_param[3 * (int64_t)g];
            G.p[1] = gate_param[3 * (int64_t)g + 1];
            G.p[2] = gate_param[3 * (int64_t)g + 2];
            G.observable = false;
            DMV_CHECK_ARG(G.kind >= DMV_GATE_H && G.kind <= DMV_GATE_CZ, "batch circuit %d: bad gate kind", i);
            DMV_CHECK_ARG(G.target >= 0 && G.target < ni, "batch circuit %d: target out of range", i);
            const bool ctl = G.kind == DMV_GATE_CNOT || G.kind == DMV_GATE_CZ;
            DMV_CHECK_ARG(!ctl || (G.control >= 0 && G.control < ni && G.control != G.target),
                          "batch circuit %d: bad control", i);
            if (!ctl)
                G.control = -1;
            gates.push_back(G);
        }
        for (const Gate &G : gates)
            if (gate_is_parametric(G.kind))
                B.np++;
        std::vector<HostOp> f = merge_ops(forward_ops(gates));
        std::vector<HostOp> a = merge_ops(adjoint_ops(gates));
        B.cost = (double)(f.size() + a.size()) * std::ldexp(1.0, ni);
        B.small = ni <= small_max;
        if (B.small) {
            B.fo = global_dev_ops(f);
            B.ao = global_dev_ops(a);
            return DMV_OK;
        }
        PlanParams pp = pp0;
        pp.n = ni;
        DMV_TRY(build_plan(f, pp, &B.pf));
        DMV_TRY(build_plan(a, pp, &B.pa));
        for (size_t p = 0; p + 1 < B.pa.ip_begin.size(); p++)
            DMV_CHECK_ARG(B.pa.ip_begin[p + 1] - B.pa.ip_begin[p] <= kBatchMaxIp,
                          "batch circuit %d: %u parameters in one reverse pass (at most %u)", i,
                          B.pa.ip_begin[p + 1] - B.pa.ip_begin[p], kBatchMaxIp);
        return DMV_OK;
    };
    {
        const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
        const int nthreads = std::max(1, std::min({hw, 16, (count + 31) / 32}));
        std::atomic<int> next{0};
        auto worker = [&]() {
            for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
                built[i].rc = build_one(i, built[i]);
                if (built[i].rc != DMV_OK)
                    built[i].err = last_error();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; t++)
            pool.emplace_back(worker);
        worker();
        for (std::thread &t : pool)
            t.join();
    }

    std::vector<BatchCircuit> circ(count);
    std::vector<DevPass> passes;
    std::vector<DevPhase> phases;
    std::vector<DevOp> ops, sops;
    std::vector<double> cost(count);
    int nmax = 1;
    int64_t total_params = 0;
    for (int i = 0; i < count; i++) {
        Built &B = built[i];
        if (B.rc != DMV_OK) {
            set_error("%s", B.err.c_str());
            return B.rc;
        }
        const int ni = n[i];
        nmax = std::max(nmax, ni);
        if (param_offset[i] != total_params) {
            set_error("param_offset[%d] is %d; it must be the number of parameters of the circuits before it (%lld)", i,
                      param_offset[i], (long long)total_params);
            return DMV_ERR_ARG;
        }
        total_params += B.np;
        BatchCircuit &C = circ[i];
        C.n = (uint32_t)ni;
        C.coef = (uint32_t)coef_offset[i];
        C.grad = (uint32_t)param_offset[i];
        cost[i] = B.cost;
        if (B.small) {
            C.small = 1;
            C.fwd_begin = (uint32_t)sops.size();
            sops.insert(sops.end(), B.fo.begin(), B.fo.end());
            C.fwd_end = (uint32_t)sops.size();
            C.adj_begin = (uint32_t)sops.size();
            sops.insert(sops.end(), B.ao.begin(), B.ao.end());
            C.adj_end = (uint32_t)sops.size();
        } else {
            C.small = 0;
            for (int which = 0; which < 2; which++) {
                const Plan &plan = which == 0 ? B.pf : B.pa;
                const uint32_t pbase = (uint32_t)passes.size();
                const uint32_t phbase = (uint32_t)phases.size();
                const uint32_t obase = (uint32_t)ops.size();
                for (DevPass P : plan.passes) {
                    P.phase_begin += phbase;
                    P.phase_end += phbase;
                    passes.push_back(P);
                }
                for (DevPhase D : plan.phases) {
                    D.op_begin += obase;
                    D.op_end += obase;
                    phases.push_back(D);
                }
                ops.insert(ops.end(), plan.ops.begin(), plan.ops.end());
                if (which == 0) {
                    C.fwd_begin = pbase;
                    C.fwd_end = (uint32_t)passes.size();
                } else {
                    C.adj_begin = pbase;
                    C.adj_end = (uint32_t)passes.size();
                }
            }
        }
        B = Built();  // release the per-circuit copies early
    }
    std::vector<uint32_t> order(count);
    for (int i = 0; i < count; i++)
        order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });

    // host work and argument checks are done: only now touch the device
    DMV_TRY(ensure_device());
    uint32_t grid = (uint32_t)std::min<int64_t>(count, (int64_t)device_cus() * per_cu);
    const uint64_t slot_amps = uint64_t(1) << std::max(nmax, K);
    const uint64_t scratch_bytes = (uint64_t)grid * 2 * slot_amps * amp_bytes(dtype);

    int64_t coef_len = 0;
    for (int i = 0; i < count; i++)
        coef_len = std::max<int64_t>(coef_len, (int64_t)coef_offset[i] + 2 * n[i] - 1);

    // device buffers
    struct Buf {
        void *p = nullptr;
        ~Buf() { release(p); }
    } b_circ, b_order, b_counter, b_passes, b_phases, b_ops, b_sops, b_coef, b_scratch, b_energy, b_grad;
    DMV_TRY(alloc(&b_circ.p, circ.size() * sizeof(BatchCircuit)));
    DMV_TRY(alloc(&b_order.p, order.size() * sizeof(uint32_t)));
    DMV_TRY(alloc(&b_counter.p, 64));
    DMV_TRY(alloc(&b_passes.p, std::max<size_t>(1, passes.size()) * sizeof(DevPass)));
    DMV_TRY(alloc(&b_phases.p, std::max<size_t>(1, phases.size()) * sizeof(DevPhase)));
    DMV_TRY(alloc(&b_ops.p, std::max<size_t>(1, ops.size()) * sizeof(DevOp)));
    DMV_TRY(alloc(&b_sops.p, std::max<size_t>(1, sops.size()) * sizeof(DevOp)));
    DMV_TRY(alloc(&b_coef.p, (uint64_t)std::max<int64_t>(1, coef_len) * sizeof(double)));
    DMV_TRY(alloc(&b_scratch.p, scratch_bytes));
    DMV_TRY(alloc(&b_energy.p, (uint64_t)count * sizeof(double)));
    DMV_TRY(alloc(&b_grad.p, (uint64_t)std::max<int64_t>(1, total_params) * sizeof(double)));
    DMV_TRY(copy_h2d(b_circ.p, circ.data(), circ.size() * sizeof(BatchCircuit)));
    DMV_TRY(copy_h2d(b_order.p, order.data(), order.size() * sizeof(uint32_t)));
    DMV_TRY(copy_h2d(b_passes.p, passes.data(), passes.size() * sizeof(DevPass)));
    DMV_TRY(copy_h2d(b_phases.p, phases.data(), phases.size() * sizeof(DevPhase)));
    DMV_TRY(copy_h2d(b_ops.p, ops.data(), ops.size() * sizeof(DevOp)));
    DMV_TRY(copy_h2d(b_sops.p, sops.data(), sops.size() * sizeof(DevOp)));
    DMV_TRY(copy_h2d(b_coef.p, coef, (uint64_t)coef_len * sizeof(double)));

    BatchLaunch L;
    L.circ = (const BatchCircuit *)b_circ.p;
    L.order = (const uint32_t *)b_order.p;
    L.count = (uint32_t)count;
    L.counter = (uint32_t *)b_counter.p;
    L.passes = (const DevPass *)b_passes.p;
    L.phases = (const DevPhase *)b_phases.p;
    L.ops = (const DevOp *)b_ops.p;
    L.sops = (const DevOp *)b_sops.p;
    L.coef = (const double *)b_coef.p;
    L.scratch = b_scratch.p;
    L.slot_amps = slot_amps;
    L.energy = (double *)b_energy.p;
    L.grad = (double *)b_grad.p;
    L.grid = grid;
    L.rlog = rlog;
    L.tlog = tlog;
    // streaming loads and stores once the scratch states take twice the
    // last-level cache (512 MiB): measured on the 4096-circuit 12-18 qubit
    // batch (6 GiB of scratch), complex128 169 -> 159 ms and complex64
    // 120-123 -> 111-118 ms
    L.nt = tune_value("nontemporal", scratch_bytes >= (uint64_t(512) << 20) ? 1 : 0) != 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        if (e0)
            (void)hipEventDestroy(e0);
        set_error("hipEventCreate failed");
        return DMV_ERR_HIP;
    }
    if (kernel_ms)
        (void)hipEventRecord(e0, stream());
    prof_begin(DMV_KC_BATCH);
    int rc = launch_batch(dtype, L);
    if (kernel_ms) {
        (void)hipEventRecord(e1, stream());
        float ms = 0.f;
        if (rc == DMV_OK && hipEventSynchronize(e1) == hipSuccess)
            (void)hipEventElapsedTime(&ms, e0, e1);
        *kernel_ms = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    DMV_TRY(rc);
    double bytes = 0.0;
    for (int i = 0; i < count; i++) {
        const double sb = std::ldexp((double)amp_bytes(dtype), n[i]);
        const BatchCircuit &C = circ[i];
        if (C.small)
            bytes += 4.0 * sb;  // reads and writes of psi and lambda stay in LDS
        else
            bytes += 2.0 * sb * (C.fwd_end - C.fwd_begin) + 4.0 * sb * (C.adj_end - C.adj_begin) + 3.0 * sb;
    }
    prof_end(DMV_KC_BATCH, bytes);
    DMV_TRY(copy_d2h(energy, b_energy.p, (uint64_t)count * sizeof(double)));
    if (total_params > 0)
        DMV_TRY(copy_d2h(grad, b_grad.p, (uint64_t)total_params * sizeof(double)));
    return DMV_OK;
}

}  // namespace dmv
