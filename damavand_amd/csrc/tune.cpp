// Launch-shape autotuning.  The reference picks grid and block sizes with an
// occupancy query before every launch (damavand-gpu/utils.cuh:62-133); here
// the shapes are measured once per (GPU architecture, qubits, dtype, kind)
// by timing the candidates on the caller's workload, kept in a process-wide
// table and in an optional JSON-lines cache file (DMV_TUNE_CACHE), and
// applied per circuit:
//  * forward: tile qubits, threads per workgroup, GEN budget per pass and
//    streaming stores of the fused passes;
//  * adjoint: tile qubits, threads and GEN budget of the reverse sweep;
//  * batch: tile shape and workgroups per CU of the persistent kernel.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace dmv {

int batch_forward_adjoint(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                          const int *gate_control, const int *gate_target, const double *gate_param,
                          const int *coef_offset, const double *coef, const int *param_offset, double *energy,
                          double *grad, const TuneChoice *pin, double *kernel_ms);

// ------------------------------------------------------------------ table --
struct TuneTable {
    std::mutex mu;
    bool loaded = false;
    std::map<std::string, TuneChoice> entries;
};

static TuneTable &table()
{
    static TuneTable *t = new TuneTable();  // never destroyed: outlives static destructors
    return *t;
}

// The number after "name": in a cache line, or `dflt`.
static double json_number(const std::string &line, const char *name, double dflt)
{
    const std::string key = std::string("\"") + name + "\"";
    size_t p = line.find(key);
    if (p == std::string::npos)
        return dflt;
    p = line.find(':', p + key.size());
    if (p == std::string::npos)
        return dflt;
    char *end = nullptr;
    const double v = strtod(line.c_str() + p + 1, &end);
    return end == line.c_str() + p + 1 ? dflt : v;
}

static std::string json_string(const std::string &line, const char *name)
{
    const std::string key = std::string("\"") + name + "\"";
    size_t p = line.find(key);
    if (p == std::string::npos)
        return "";
    p = line.find('"', line.find(':', p + key.size()));
    if (p == std::string::npos)
        return "";
    const size_t q = line.find('"', p + 1);
    return q == std::string::npos ? "" : line.substr(p + 1, q - p - 1);
}

// Parses one cache line; false when it is not an entry.
bool parse_tune_line(const std::string &line, std::string *key, TuneChoice *c)
{
    *key = json_string(line, "key");
    if (key->empty())
        return false;
    c->k = (int)json_number(line, "tile_qubits", 0);
    c->tlog = (int)json_number(line, "threads_log2", 0);
    c->max_gen = (int)json_number(line, "max_gen_per_pass", 0);
    c->nt = (int)json_number(line, "nontemporal", -1);
    c->wgs = (int)json_number(line, "wgs_per_cu", 0);
    c->ms = json_number(line, "ms", 0.0);
    return c->k > 0 && c->k <= kMaxTileQubits && c->tlog >= 0 && c->tlog <= kMaxThreadBits;
}

static void load_locked(TuneTable &T)
{
    if (T.loaded)
        return;
    T.loaded = true;
    const char *path = getenv("DMV_TUNE_CACHE");
    if (!path || !*path)
        return;
    FILE *f = fopen(path, "r");
    if (!f)
        return;
    char buf[1024];
    while (fgets(buf, sizeof(buf), f)) {
        std::string key;
        TuneChoice c;
        if (parse_tune_line(buf, &key, &c))
            T.entries[key] = c;  // later lines win
    }
    fclose(f);
}

static std::string arch_name()
{
    static std::string name = [] {
        hipDeviceProp_t prop;
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            std::string a = prop.gcnArchName;
            return a.substr(0, a.find(':'));
        }
        (void)hipGetLastError();
        return std::string("gfx950");
    }();
    return name;
}

std::string tune_key(int n, int dtype, const char *kind)
{
    return arch_name() + "/n" + std::to_string(n) + (dtype == DMV_C64 ? "/c64/" : "/c128/") + kind;
}

bool tune_lookup(const std::string &key, TuneChoice *out)
{
    TuneTable &T = table();
    std::lock_guard<std::mutex> lk(T.mu);
    load_locked(T);
    auto it = T.entries.find(key);
    if (it == T.entries.end())
        return false;
    *out = it->second;
    return true;
}

void tune_store(const std::string &key, const TuneChoice &c)
{
    TuneTable &T = table();
    std::lock_guard<std::mutex> lk(T.mu);
    load_locked(T);
    T.entries[key] = c;
    const char *path = getenv("DMV_TUNE_CACHE");
    if (!path || !*path)
        return;
    if (FILE *f = fopen(path, "a")) {
        fprintf(f,
                "{\"key\": \"%s\", \"tile_qubits\": %d, \"threads_log2\": %d, \"max_gen_per_pass\": %d, "
                "\"nontemporal\": %d, \"wgs_per_cu\": %d, \"ms\": %.6f}\n",
                key.c_str(), c.k, c.tlog, c.max_gen, c.nt, c.wgs, c.ms);
        fclose(f);
    }
}

// Forgets the table so that the next lookup reads DMV_TUNE_CACHE again.
void tune_table_reset()
{
    TuneTable &T = table();
    std::lock_guard<std::mutex> lk(T.mu);
    T.entries.clear();
    T.loaded = false;
}

// -------------------------------------------------------------- autotuners --
namespace {

// Events and a scratch state that are released on every exit path.
struct Timer {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Timer()
    {
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
    }
    int init()
    {
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            set_error("hipEventCreate failed");
            return DMV_ERR_HIP;
        }
        return DMV_OK;
    }
};

struct Scratch {
    dmv_state *s = nullptr;
    ~Scratch() { state_destroy(s); }
};

struct Plan1 {
    DevicePlan dp;
    ~Plan1() { free_plan(dp); }
};

// Candidate tile shapes (tile qubits, log2 threads) with kernels in this
// library, largest first.
std::vector<std::pair<int, int>> shapes(int dtype, bool adjoint)
{
    if (adjoint)
        return dtype == DMV_C64 ? std::vector<std::pair<int, int>>{{12, 8}, {11, 8}, {10, 6}, {9, 6}}
                                : std::vector<std::pair<int, int>>{{11, 8}, {10, 8}, {9, 6}};
    return dtype == DMV_C64 ? std::vector<std::pair<int, int>>{{13, 8}, {12, 8}, {11, 8}, {10, 6}, {9, 6}}
                            : std::vector<std::pair<int, int>>{{12, 8}, {12, 9}, {11, 8}, {11, 7}, {10, 8}, {9, 6}};
}

// Times `run` on the library stream: the smaller of two runs after a warm-up.
template <class F> int time_runs(Timer &T, F run, float *ms)
{
    DMV_TRY(run());
    float best = 1e30f;
    for (int r = 0; r < 2; r++) {
        (void)hipEventRecord(T.e0, stream());
        DMV_TRY(run());
        (void)hipEventRecord(T.e1, stream());
        if (hipEventSynchronize(T.e1) != hipSuccess) {
            set_error("hipEventSynchronize failed");
            return DMV_ERR_HIP;
        }
        float t = 0.f;
        (void)hipEventElapsedTime(&t, T.e0, T.e1);
        best = std::min(best, t);
    }
    *ms = best;
    return DMV_OK;
}

}  // namespace

int autotune_forward(dmv_circuit *c, dmv_state *s, int *chosen)
{
    DMV_CHECK_ARG(c->n == s->n, "circuit has %d qubits but the state has %d", c->n, s->n);
    DMV_TRY(ensure_device());
    const int small_cap = s->dtype == DMV_C64 ? kSmallMaxQubitsC64 : kSmallMaxQubitsC128;
    if (s->n <= small_cap) {  // one whole-state launch: nothing to choose
        if (chosen)
            *chosen = s->n;
        return DMV_OK;
    }
    Timer T;
    DMV_TRY(T.init());
    Scratch sc;  // the candidates run on a scratch state: the caller's state is left as it was
    DMV_TRY(state_create(s->n, s->dtype, &sc.s));
    const std::vector<HostOp> ops = forward_ops(c->gates);
    TuneChoice best;
    float best_ms = 1e30f;
    for (auto [k, tlog] : shapes(s->dtype, false)) {
        if (k + 3 > s->n)
            continue;
        for (int gen : {6, 8, 12, 16}) {
            TuneChoice cand;
            cand.k = k;
            cand.tlog = tlog;
            cand.max_gen = gen;
            cand.nt = 0;
            Plan1 P;
            float ms = 0.f;
            DMV_TRY(prepare_plan(P.dp, c->version, c->n, sc.s->dtype, false, ops, sc.s->layout, &cand));
            DMV_TRY(time_runs(T, [&] { return run_forward_plan(P.dp, sc.s); }, &ms));
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
    }
    if (!best.set()) {
        if (chosen)
            *chosen = 0;
        return DMV_OK;
    }
    {  // streaming loads and stores on the best shape
        TuneChoice cand = best;
        cand.nt = 1;
        Plan1 P;
        float ms = 0.f;
        DMV_TRY(prepare_plan(P.dp, c->version, c->n, sc.s->dtype, false, ops, sc.s->layout, &cand));
        DMV_TRY(time_runs(T, [&] { return run_forward_plan(P.dp, sc.s); }, &ms));
        if (ms < best_ms) {
            best_ms = ms;
            best = cand;
        }
    }
    best.ms = best_ms;
    c->tune_fwd = best;
    tune_store(tune_key(s->n, s->dtype, "forward"), best);
    if (chosen)
        *chosen = best.k;
    return DMV_OK;
}

int autotune_adjoint(dmv_circuit *c, int dtype, int *chosen)
{
    DMV_CHECK_ARG(dtype == DMV_C128 || dtype == DMV_C64, "dtype must be DMV_C128 or DMV_C64");
    DMV_TRY(ensure_device());
    const int small_cap = (dtype == DMV_C64 ? kSmallMaxQubitsC64 : kSmallMaxQubitsC128) - 1;
    if (c->n <= small_cap) {
        if (chosen)
            *chosen = c->n;
        return DMV_OK;
    }
    Timer T;
    DMV_TRY(T.init());
    Scratch psi, lam;
    DMV_TRY(state_create(c->n, dtype, &psi.s));
    DMV_TRY(state_create(c->n, dtype, &lam.s));
    const std::vector<double> a(c->n, 1.0);  // H = sum_q Z_q: the sweep's cost does not depend on H
    Observable obs;
    obs.a = a.data();
    TuneChoice best;
    float best_ms = 1e30f;
    for (auto [k, tlog] : shapes(dtype, true)) {
        if (k + 3 > c->n)
            continue;
        for (int gen : {8, 16}) {
            TuneChoice cand;
            cand.k = k;
            cand.tlog = tlog;
            cand.max_gen = gen;
            Plan1 P;
            float ms = 0.f;
            DMV_TRY(time_runs(T,
                              [&] {
                                  return circuit_adjoint(c, psi.s, lam.s, obs, nullptr, nullptr, P.dp, &cand);
                              },
                              &ms));
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
    }
    if (!best.set()) {
        if (chosen)
            *chosen = 0;
        return DMV_OK;
    }
    best.ms = best_ms;
    c->tune_adj = best;
    tune_store(tune_key(c->n, dtype, "adjoint"), best);
    if (chosen)
        *chosen = best.k;
    return DMV_OK;
}

int autotune_batch(int count, int dtype, const int *n, const int *gate_offset, const int *gate_kind,
                   const int *gate_control, const int *gate_target, const double *gate_param, const int *coef_offset,
                   const double *coef, const int *param_offset, int *chosen)
{
    DMV_CHECK_ARG(count >= 1 && n, "empty batch");
    int nmax = 0;
    int64_t n_params = 0;
    for (int i = 0; i < count; i++)
        nmax = std::max(nmax, n[i]);
    for (int i = 0; i < count; i++)
        for (int g = gate_offset[i]; g < gate_offset[i + 1]; g++)
            n_params += gate_kind[g] == DMV_GATE_RX || gate_kind[g] == DMV_GATE_RY || gate_kind[g] == DMV_GATE_RZ;
    std::vector<double> energy(count), grad(std::max<int64_t>(1, n_params));
    TuneChoice best;
    double best_ms = 1e300;
    for (auto [k, tlog] : std::vector<std::pair<int, int>>{{11, 8}, {10, 8}, {10, 7}, {12, 9}})
        for (int wgs : {1, 2, 3, 4}) {
            TuneChoice cand;
            cand.k = k;
            cand.tlog = tlog;
            cand.wgs = wgs;
            double ms = 0.0, t = 0.0;
            for (int r = 0; r < 2; r++) {  // the second run is timed
                DMV_TRY(batch_forward_adjoint(count, dtype, n, gate_offset, gate_kind, gate_control, gate_target,
                                              gate_param, coef_offset, coef, param_offset, energy.data(),
                                              grad.data(), &cand, &t));
                ms = t;
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
    best.ms = best_ms;
    tune_store(tune_key(nmax, dtype, "batch"), best);
    if (chosen)
        *chosen = best.k;
    return DMV_OK;
}

}  // namespace dmv
