// Device context, profiling, memory, states and circuit execution.
#include "runtime.hpp"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "device.hpp"
#include "kernels.hpp"

namespace dmv {

// ------------------------------------------------------------------ errors --
static thread_local std::string g_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

const char *last_error() { return g_error.c_str(); }

// ----------------------------------------------------------------- context --
struct ProfRec {
    hipEvent_t a, b;
    int kc;
    double bytes;
};

// Process-wide library state.  Calls on different handles may come from
// different threads: initialisation is double-checked on an atomic, and the
// profiling records and the tuning table each have their own lock.  All work
// goes to one library stream (a thread capturing a hipGraph launches on its
// own capture stream instead, see stream()).
struct Context {
    std::mutex mu;
    std::atomic<bool> init{false};
    int device = 0;
    hipStream_t stream = nullptr;
    std::atomic<bool> prof{false};
    std::mutex prof_mu;  // pending, free_events and the totals
    std::vector<ProfRec> pending;
    std::vector<hipEvent_t> free_events;
    int64_t launches[DMV_KC_COUNT] = {};
    double ms[DMV_KC_COUNT] = {};
    double bytes[DMV_KC_COUNT] = {};
    std::mutex tune_mu;  // tune
    std::map<std::string, int64_t> tune;
    std::atomic<uint64_t> tune_epoch{0};  // bumped by every tune_set: cached plans rebuild
};

// Per thread: the event a prof_begin opened, whether profiling is held off
// (inside a graph capture), and the stream a capture launches on.
static thread_local hipEvent_t t_open = nullptr;
static thread_local bool t_prof_off = false;
static thread_local hipStream_t t_capture = nullptr;

static Context &ctx()
{
    static Context *c = new Context();  // never destroyed: outlives static destructors
    return *c;
}

int ensure_device()
{
    Context &C = ctx();
    if (!C.init.load(std::memory_order_acquire)) {
        std::lock_guard<std::mutex> lk(C.mu);
        if (!C.init.load(std::memory_order_relaxed)) {
            int count = 0;
            hipError_t e = hipGetDeviceCount(&count);
            if (e != hipSuccess || count <= 0) {
                set_error("no GPU visible to HIP (%s)", e == hipSuccess ? "0 devices" : hipGetErrorString(e));
                return DMV_ERR_NO_DEVICE;
            }
            C.device = 0;
            if ((e = hipSetDevice(C.device)) != hipSuccess) {
                set_error("hipSetDevice: %s", hipGetErrorString(e));
                return DMV_ERR_HIP;
            }
            // default priority, non-blocking with respect to the null stream
            if ((e = hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking)) != hipSuccess) {
                set_error("hipStreamCreate: %s", hipGetErrorString(e));
                return DMV_ERR_HIP;
            }
            C.init.store(true, std::memory_order_release);
        }
    }
    hipError_t e = hipSetDevice(C.device);
    if (e != hipSuccess) {
        set_error("hipSetDevice: %s", hipGetErrorString(e));
        return DMV_ERR_HIP;
    }
    return DMV_OK;
}

hipStream_t stream() { return t_capture ? t_capture : ctx().stream; }

int device_cus()
{
    static std::atomic<int> cus{0};
    int c = cus.load(std::memory_order_relaxed);
    if (c == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, ctx().device) == hipSuccess && prop.multiProcessorCount > 0)
            c = prop.multiProcessorCount;
        else
            c = 256;
        cus.store(c, std::memory_order_relaxed);
    }
    return c;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s launch failed: %s", what, hipGetErrorString(e));
        return DMV_ERR_HIP;
    }
    return DMV_OK;
}

#define HIPCALL(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            set_error("%s: %s", #call, hipGetErrorString(e_));                                     \
            return DMV_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)

int synchronize()
{
    DMV_TRY(ensure_device());
    HIPCALL(hipStreamSynchronize(stream()));
    return DMV_OK;
}

int alloc(void **p, uint64_t bytes)
{
    *p = nullptr;
    if (bytes == 0)
        return DMV_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_error("device allocation of %llu bytes failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
        return DMV_ERR_MEMORY;
    }
    return DMV_OK;
}

void release(void *p)
{
    if (p) {
        (void)hipStreamSynchronize(stream());
        (void)hipFree(p);
    }
}

int copy_h2d(void *dst, const void *src, uint64_t bytes)
{
    if (bytes == 0)
        return DMV_OK;
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    HIPCALL(hipStreamSynchronize(stream()));
    return DMV_OK;
}

int copy_d2h(void *dst, const void *src, uint64_t bytes)
{
    if (bytes == 0)
        return DMV_OK;
    HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    HIPCALL(hipStreamSynchronize(stream()));
    return DMV_OK;
}

// --------------------------------------------------------------- profiling --
// get_event and collect run with prof_mu held.
static hipEvent_t get_event()
{
    Context &C = ctx();
    if (!C.free_events.empty()) {
        hipEvent_t e = C.free_events.back();
        C.free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

static void collect()
{
    Context &C = ctx();
    if (C.pending.empty())
        return;
    (void)hipStreamSynchronize(C.stream);  // every record is on the library stream
    for (ProfRec &r : C.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            C.launches[r.kc]++;
            C.ms[r.kc] += ms;
            C.bytes[r.kc] += r.bytes;
        }
        C.free_events.push_back(r.a);
        C.free_events.push_back(r.b);
    }
    C.pending.clear();
}

void prof_begin(int)
{
    Context &C = ctx();
    if (!C.prof.load(std::memory_order_relaxed) || t_prof_off)
        return;
    std::lock_guard<std::mutex> lk(C.prof_mu);
    t_open = get_event();
    if (t_open)
        (void)hipEventRecord(t_open, C.stream);
}

void prof_end(int kc, double bytes)
{
    Context &C = ctx();
    if (!C.prof.load(std::memory_order_relaxed) || t_prof_off || !t_open)
        return;
    std::lock_guard<std::mutex> lk(C.prof_mu);
    hipEvent_t b = get_event();
    if (!b)
        return;
    (void)hipEventRecord(b, C.stream);
    C.pending.push_back({t_open, b, kc, bytes});
    t_open = nullptr;
    if (C.pending.size() >= 4096)
        collect();
}

int64_t tune_value(const char *name, int64_t dflt)
{
    Context &C = ctx();
    {
        std::lock_guard<std::mutex> lk(C.tune_mu);
        auto it = C.tune.find(name);
        if (it != C.tune.end())
            return it->second;
    }
    const std::string env = std::string("DMV_") + name;
    std::string up;
    for (char ch : env)
        up += (char)toupper((unsigned char)ch);
    if (const char *v = getenv(up.c_str()))
        return atoll(v);
    return dflt;
}


// Layout of a new state.  The span index is folded on states of 2^15 spans
// and more (32 qubits in complex128): there, tiles over the highest qubits
// otherwise run at 2.4 TB/s (measured: 33 qubits, qubits 23-31), while at 30
// qubits the fold measured about 1% slower on the layered workload.
// Tunables: page_fold_span_bits (span-index width from which to fold),
// page_fold (fold width).
StateLayout default_layout(int n, int dtype)
{
    const bool c64 = dtype == DMV_C64;
    const StateLayout plain = state_layout(n, c64, 0);
    const int64_t from = tune_value("page_fold_span_bits", 15);
    const int64_t width = tune_value("page_fold", kFoldBase);
    if (!plain.shift || n - (int)plain.shift < from || width <= 0)
        return plain;
    return state_layout(n, c64, (uint32_t)width);
}

// Single-gate kernels: one pair per thread; nontemporal loads/stores unless
// the target is one of the lowest qubits, where the two streams share lines
// (measured best on MI355X at 30 qubits).
int gate_pairs_per_thread() { return (int)tune_value("gate_pairs_per_thread", 1); }

bool gate_nontemporal(uint32_t target, int dtype)
{
    const int64_t v = tune_value("gate_nontemporal", -1);
    if (v >= 0)
        return v != 0;
    return target >= (uint32_t)coalesce_bits(dtype);
}

// ------------------------------------------------------------------- plans --
PlanParams tile_params(int n, int dtype, bool adjoint, StateLayout layout, const TuneChoice *pin)
{
    PlanParams pp;
    pp.n = n;
    pp.layout = layout;
    pp.low_bits = coalesce_bits(dtype);
    // wider contiguous runs per lane group (tunable; fewer free tile qubits)
    const int64_t runs = tune_value("coalesce_qubits", 0);
    if (runs > pp.low_bits && runs <= 6)
        pp.low_bits = (int)runs;
    pp.swz_bits = dtype == DMV_C64 ? 4 : 3;
    // defaults: 256 threads; k = 12 (complex128) / 13 (complex64) forward,
    // 11 for reverse sweeps, which hold psi and lambda in registers together
    int k = dtype == DMV_C64 ? (adjoint ? 11 : 13) : (adjoint ? 11 : 12);
    // streaming (nontemporal) loads and stores: forward passes from 2 GiB
    // states up (measured 1.5-3.5% faster per pass at 27, 30 and 33 qubits,
    // complex128, and slower at 24 qubits, where much of the state stays in
    // the last-level cache); reverse sweeps once psi and lambda together take
    // 512 MiB, twice that cache (measured 1% faster per pass at 24 qubits,
    // 1-3% at 27 and 3.5% at 30)
    const uint64_t state_bytes = (uint64_t)amp_bytes(dtype) << n;
    int tlog = 8, gen = 8;
    int nt = adjoint ? (2 * state_bytes >= (uint64_t(512) << 20) ? 1 : 0) : (state_bytes >= (uint64_t(2) << 30) ? 1 : 0);
    bool chosen = false;  // the shape was measured for this size (table or circuit)
    bool gen_set = false;  // the GEN budget was given (table, tunable, circuit)
    TuneChoice t;
    if (tune_lookup(tune_key(n, dtype, adjoint ? "adjoint" : "forward"), &t) && t.set()) {
        k = t.k;
        tlog = t.tlog > 0 ? t.tlog : tlog;
        gen = t.max_gen > 0 ? t.max_gen : gen;
        gen_set = t.max_gen > 0;
        nt = t.nt >= 0 ? t.nt : nt;
        chosen = true;
    }
    const int64_t k_set = tune_value(adjoint ? "adjoint_tile_qubits" : "tile_qubits", 0);
    const int64_t t_set = tune_value("tile_threads_log2", 0);
    if (k_set > 0 || t_set > 0) {
        k = k_set > 0 ? (int)k_set : k;
        tlog = t_set > 0 ? (int)t_set : 8;
        chosen = false;
    }
    if (tune_value("max_gen_per_pass", -1) > 0) {
        gen = (int)tune_value("max_gen_per_pass", gen);
        gen_set = true;
    }
    nt = (int)tune_value("nontemporal", nt);
    if (pin && pin->set()) {
        k = pin->k;
        tlog = pin->tlog > 0 ? pin->tlog : tlog;
        gen = pin->max_gen > 0 ? pin->max_gen : gen;
        gen_set = gen_set || pin->max_gen > 0;
        nt = pin->nt >= 0 ? pin->nt : nt;
        chosen = true;
    }
    // small states: fewer threads and smaller tiles keep more workgroups busy
    if (!chosen && n < k + 6) {
        tlog = 6;
        k = adjoint ? 9 : (dtype == DMV_C64 ? 10 : 9);
    }
    if (k > n)
        k = n;
    pp.k = k;
    pp.tlog = tlog;
    pp.rlog = k - tlog;
    pp.max_gen_per_pass = gen;
    // forward plans choose 8 or 12 per circuit unless a budget was given
    pp.gen_auto = !adjoint && !gen_set ? 1 : 0;
    pp.nontemporal = nt != 0 ? 1 : 0;
    // forward passes split their transposes where the shape has half-LDS kernels
    pp.half_lds = !adjoint && tune_value("half_lds", 1) != 0 && tile_half_exists(dtype, pp.rlog, pp.tlog) ? 1 : 0;
    return pp;
}

void free_plan(DevicePlan &dp)
{
    if (dp.graph) {
        (void)hipGraphExecDestroy(dp.graph);
        dp.graph = nullptr;
    }
    release(dp.d_passes);
    release(dp.d_phases);
    release(dp.d_ops);
    release(dp.d_ip_slots);
    dp.d_passes = nullptr;
    dp.d_phases = nullptr;
    dp.d_ops = nullptr;
    dp.d_ip_slots = nullptr;
    dp.cap_passes = dp.cap_phases = dp.cap_ops = dp.cap_ip = 0;
    dp.valid = false;
}

template <class T> static int upload(T *&dptr, size_t &cap, const std::vector<T> &v)
{
    if (v.size() > cap) {
        release(dptr);
        dptr = nullptr;
        const size_t want = v.size() + v.size() / 2 + 4;
        DMV_TRY(alloc((void **)&dptr, want * sizeof(T)));
        cap = want;
    }
    return copy_h2d(dptr, v.data(), v.size() * sizeof(T));
}

int prepare_plan(DevicePlan &dp, uint64_t version, int n, int dtype, bool adjoint, const std::vector<HostOp> &ops,
                 StateLayout layout, const TuneChoice *pin)
{
    const int small_cap = (dtype == DMV_C64 ? kSmallMaxQubitsC64 : kSmallMaxQubitsC128) - (adjoint ? 1 : 0);
    const bool small = n <= small_cap;
    PlanParams pp = small ? PlanParams{} : tile_params(n, dtype, adjoint, layout, pin);
    const bool same_shape = dp.valid && dp.dtype == dtype && dp.small == small &&
                            (small || (dp.pp.k == pp.k && dp.pp.tlog == pp.tlog && dp.pp.max_gen_per_pass == pp.max_gen_per_pass &&
                                       dp.pp.gen_auto == pp.gen_auto &&
                                       dp.pp.nontemporal == pp.nontemporal && dp.pp.half_lds == pp.half_lds &&
                                       same_layout(dp.pp.layout, pp.layout)));
    if (same_shape && dp.version == version && dp.epoch == tune_epoch())
        return DMV_OK;
    std::vector<HostOp> merged = merge_ops(ops);
    dp.small = small;
    dp.dtype = dtype;
    dp.pp = pp;
    if (small) {
        dp.small_ops = global_dev_ops(merged);
        dp.plan = Plan();
        for (const DevOp &o : dp.small_ops)
            if (o.slot + 1 > dp.plan.n_slots)
                dp.plan.n_slots = o.slot + 1;
        DMV_TRY(upload(dp.d_ops, dp.cap_ops, dp.small_ops));
    } else {
        Plan plan;
        DMV_TRY(build_plan(merged, pp, &plan));
        dp.plan = std::move(plan);
        DMV_TRY(upload(dp.d_passes, dp.cap_passes, dp.plan.passes));
        DMV_TRY(upload(dp.d_phases, dp.cap_phases, dp.plan.phases));
        DMV_TRY(upload(dp.d_ops, dp.cap_ops, dp.plan.ops));
        if (!dp.plan.ip_slots.empty())
            DMV_TRY(upload(dp.d_ip_slots, dp.cap_ip, dp.plan.ip_slots));
    }
    dp.version = version;
    dp.epoch = tune_epoch();
    dp.build_id++;
    dp.valid = true;
    return DMV_OK;
}

// ------------------------------------------------------------------ states --
static uint64_t ws_doubles_for(int n)
{
    const int B = n < 12 ? n : 12;
    const uint64_t nblk = uint64_t(1) << (n - B);
    return nblk * 13 + 64;
}

int state_create(int n, int dtype, dmv_state **out)
{
    DMV_CHECK_ARG(out != nullptr, "null output pointer");
    *out = nullptr;
    DMV_CHECK_ARG(n >= 1 && n <= kMaxQubits, "num_qubits must be in [1, %d], got %d", kMaxQubits, n);
    DMV_CHECK_ARG(dtype == DMV_C128 || dtype == DMV_C64, "dtype must be DMV_C128 or DMV_C64");
    DMV_TRY(ensure_device());
    dmv_state *s = new dmv_state();
    s->n = n;
    s->dtype = dtype;
    s->bytes = (uint64_t(1) << n) * (uint64_t)amp_bytes(dtype);
    s->alloc_bytes = s->bytes;  // the swizzled layout is a permutation within the plain size
    s->layout = default_layout(n, dtype);
    s->ws_doubles = ws_doubles_for(n);
    const uint64_t nblk = s->ws_doubles;
    int rc = DMV_ERR_MEMORY;
    if (tune_value("spread_states", 1) && spread_alloc(s->alloc_bytes, &s->spread) == DMV_OK)
        s->amp = s->spread.va, rc = DMV_OK;
    if (rc != DMV_OK)
        rc = alloc(&s->amp, s->alloc_bytes);
    if (rc == DMV_OK)
        rc = alloc((void **)&s->ws, s->ws_doubles * sizeof(double));
    if (rc == DMV_OK)
        rc = alloc((void **)&s->cum, nblk * sizeof(double));
    if (rc == DMV_OK)
        rc = alloc((void **)&s->scalars, 64 * sizeof(double));
    if (rc == DMV_OK)
        rc = state_reset(s, 0, false);
    if (rc != DMV_OK) {
        state_destroy(s);
        return rc;
    }
    *out = s;
    return DMV_OK;
}

void state_destroy(dmv_state *s)
{
    if (!s)
        return;
    if (s->spread.va)
        spread_free(&s->spread);
    else
        release(s->amp);
    release(s->ws);
    release(s->cum);
    release(s->scalars);
    release(s->stage);
    release(s->terms);
    if (s->bounce) {
        (void)hipStreamSynchronize(stream());
        (void)hipHostFree(s->bounce);
    }
    for (hipEvent_t &e : s->bounce_ev)
        if (e)
            (void)hipEventDestroy(e);
    delete s;
}

int state_reset(dmv_state *s, uint64_t basis, bool zero)
{
    DMV_TRY(ensure_device());
    prof_begin(DMV_KC_REDUCE);
    DMV_TRY(launch_init_basis(s->dtype, s->amp, s->n, s->layout, zero ? ~uint64_t(0) : basis));
    prof_end(DMV_KC_REDUCE, (double)s->bytes);
    return DMV_OK;
}

static const uint64_t kChunk = uint64_t(1) << 24;  // amplitudes per staged transfer

// The state's own staging buffer for split re/im chunks (allocated on first
// use, so calls on different states never share one).
static int staging(dmv_state *s, double **re, double **im, uint64_t count)
{
    if (count > s->stage_cap) {
        release(s->stage);
        s->stage = nullptr;
        s->stage_cap = 0;
        DMV_TRY(alloc((void **)&s->stage, 2 * count * sizeof(double)));
        s->stage_cap = count;
    }
    *re = s->stage;
    *im = s->stage + count;
    return DMV_OK;
}

// Large device-to-host copies of a state's results go through its own
// pinned buffer, in two halves so that the DMA of one piece overlaps the host
// copy of the other: the transfer then runs at link speed whatever the
// caller's memory (a fresh, never touched or never pinned array made the
// pageable path take about 10 ms for 8 MiB on first use).
static constexpr uint64_t kBounceHalf = uint64_t(8) << 20;

static int bounce_d2h(dmv_state *s, void *dst, const void *src, uint64_t bytes)
{
    if (bytes <= (uint64_t(64) << 10))
        return copy_d2h(dst, src, bytes);
    if (!s->bounce) {
        // events first, then the buffer: the state holds either all three or none
        for (hipEvent_t &e : s->bounce_ev)
            if (!e)
                HIPCALL(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        void *b = nullptr;
        HIPCALL(hipHostMalloc(&b, 2 * kBounceHalf, hipHostMallocDefault));
        s->bounce = b;
    }
    char *buf[2] = {(char *)s->bounce, (char *)s->bounce + kBounceHalf};
    const uint64_t pieces = (bytes + kBounceHalf - 1) / kBounceHalf;
    auto issue = [&](uint64_t i) -> int {
        const uint64_t off = i * kBounceHalf, len = std::min(kBounceHalf, bytes - off);
        HIPCALL(hipMemcpyAsync(buf[i & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, stream()));
        HIPCALL(hipEventRecord(s->bounce_ev[i & 1], stream()));
        return DMV_OK;
    };
    DMV_TRY(issue(0));
    for (uint64_t i = 0; i < pieces; i++) {
        // piece i + 1 goes into the other half while piece i is copied out
        if (i + 1 < pieces)
            DMV_TRY(issue(i + 1));
        HIPCALL(hipEventSynchronize(s->bounce_ev[i & 1]));
        const uint64_t off = i * kBounceHalf, len = std::min(kBounceHalf, bytes - off);
        std::memcpy((char *)dst + off, buf[i & 1], len);
    }
    return DMV_OK;
}

int state_get(dmv_state *s, double *re, double *im)
{
    DMV_TRY(ensure_device());
    const uint64_t n_amp = uint64_t(1) << s->n;
    const uint64_t chunk = n_amp < kChunk ? n_amp : kChunk;
    double *dre, *dim;
    DMV_TRY(staging(s, &dre, &dim, chunk));
    for (uint64_t b = 0; b < n_amp; b += chunk) {
        DMV_TRY(launch_split(s->dtype, s->amp, s->layout, b, chunk, dre, dim));
        if (re)
            DMV_TRY(bounce_d2h(s, re + b, dre, chunk * sizeof(double)));
        if (im)
            DMV_TRY(bounce_d2h(s, im + b, dim, chunk * sizeof(double)));
    }
    return DMV_OK;
}

int state_set(dmv_state *s, const double *re, const double *im)
{
    DMV_TRY(ensure_device());
    const uint64_t n_amp = uint64_t(1) << s->n;
    const uint64_t chunk = n_amp < kChunk ? n_amp : kChunk;
    double *dre, *dim;
    DMV_TRY(staging(s, &dre, &dim, chunk));
    for (uint64_t b = 0; b < n_amp; b += chunk) {
        DMV_TRY(copy_h2d(dre, re + b, chunk * sizeof(double)));
        DMV_TRY(copy_h2d(dim, im + b, chunk * sizeof(double)));
        DMV_TRY(launch_join(s->dtype, s->amp, s->layout, b, chunk, dre, dim));
    }
    return synchronize();
}

int state_probs(dmv_state *s, double *host)
{
    DMV_TRY(ensure_device());
    const uint64_t n_amp = uint64_t(1) << s->n;
    const uint64_t chunk = n_amp < kChunk ? n_amp : kChunk;
    double *dre, *dim;
    DMV_TRY(staging(s, &dre, &dim, chunk));
    for (uint64_t b = 0; b < n_amp; b += chunk) {
        prof_begin(DMV_KC_REDUCE);
        DMV_TRY(launch_probs(s->dtype, s->amp, s->layout, b, chunk, dre));
        prof_end(DMV_KC_REDUCE, (double)chunk * (amp_bytes(s->dtype) + 8));
        DMV_TRY(bounce_d2h(s, host + b, dre, chunk * sizeof(double)));
    }
    return DMV_OK;
}

// ---------------------------------------------------------------- circuits --
static int forward_passes(DevicePlan &dp, dmv_state *s)
{
    if (dp.small) {
        prof_begin(DMV_KC_SMALL);
        DMV_TRY(launch_small(s->dtype, s->amp, s->n, dp.d_ops, (uint32_t)dp.small_ops.size()));
        prof_end(DMV_KC_SMALL, 2.0 * (double)s->bytes);
        return DMV_OK;
    }
    const bool nt = dp.pp.nontemporal != 0;
    for (uint32_t p = 0; p < dp.plan.passes.size(); p++) {
        prof_begin(DMV_KC_TILE);
        DMV_TRY(launch_tile(s->dtype, dp.pp.rlog, dp.pp.tlog, nt, dp.plan.passes[p].lds_half != 0, s->amp, s->n,
                            s->layout, dp.d_passes, dp.d_phases, dp.d_ops, p));
        prof_end(DMV_KC_TILE, 2.0 * (double)s->bytes);
    }
    return DMV_OK;
}

static thread_local hipStream_t t_capture_stream = nullptr;

static int forward_graph(DevicePlan &dp, dmv_state *s)
{
    Context &C = ctx();
    if (!dp.graph || dp.graph_state != s->amp || dp.graph_build != dp.build_id) {
        if (dp.graph) {
            (void)hipGraphExecDestroy(dp.graph);
            dp.graph = nullptr;
        }
        // This thread's launches go to its own capture stream, so that work
        // other threads put on the library stream meanwhile stays out of the
        // graph; no event records inside the capture.
        if (!t_capture_stream)
            HIPCALL(hipStreamCreateWithFlags(&t_capture_stream, hipStreamNonBlocking));
        HIPCALL(hipStreamBeginCapture(t_capture_stream, hipStreamCaptureModeThreadLocal));
        t_capture = t_capture_stream;
        t_prof_off = true;
        int rc = forward_passes(dp, s);
        t_capture = nullptr;
        t_prof_off = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(t_capture_stream, &g);
        if (rc != DMV_OK) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
            return DMV_ERR_HIP;
        }
        e = hipGraphInstantiate(&dp.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            dp.graph = nullptr;
            set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
            return DMV_ERR_HIP;
        }
        dp.graph_state = s->amp;
        dp.graph_build = dp.build_id;
    }
    prof_begin(DMV_KC_TILE);
    HIPCALL(hipGraphLaunch(dp.graph, C.stream));
    prof_end(DMV_KC_TILE, 0.0);
    return DMV_OK;
}

int circuit_forward(dmv_circuit *c, dmv_state *s, int mode)
{
    DMV_CHECK_ARG(c && s, "null circuit or state");
    DMV_CHECK_ARG(c->n == s->n, "circuit has %d qubits but the state has %d", c->n, s->n);
    DMV_TRY(ensure_device());
    if (mode == DMV_FORWARD_PER_GATE) {
        const int ppt = gate_pairs_per_thread();
        for (const HostOp &h : forward_ops(c->gates)) {
            const bool nt = gate_nontemporal((uint32_t)h.target, s->dtype);
            DevOp op;
            std::memset(&op, 0, sizeof(op));
            op.kind = h.kind;
            op.target = (uint32_t)h.target;
            op.ctrl = h.ctrl;
            op.slot = -1;
            std::memcpy(op.m, h.m, sizeof(op.m));
            const uint64_t n_amp = uint64_t(1) << s->n;
            const int nc = popcount64(h.ctrl);
            const bool phase_only = h.kind == OP_DIAG && h.m[0] == 1.0 && h.m[1] == 0.0;
            const double touched = (double)(n_amp >> (nc + (phase_only ? 1 : 0)));
            prof_begin(DMV_KC_GATE);
            DMV_TRY(launch_gate(s->dtype, s->amp, s->n, s->layout, op, ppt, nt));
            prof_end(DMV_KC_GATE, 2.0 * touched * amp_bytes(s->dtype));
        }
        return DMV_OK;
    }
    DMV_CHECK_ARG(mode == DMV_FORWARD_FUSED, "unknown forward mode %d", mode);
    DMV_TRY(prepare_plan(c->fwd, c->version, c->n, s->dtype, false, forward_ops(c->gates), s->layout, &c->tune_fwd));
    if (tune_value("graph", 0) && !c->fwd.small)
        return forward_graph(c->fwd, s);
    return forward_passes(c->fwd, s);
}

int run_forward_plan(DevicePlan &dp, dmv_state *s) { return forward_passes(dp, s); }

// -------------------------------------------------------- profiling access --
int profile_enable(int on)
{
    DMV_TRY(ensure_device());
    std::lock_guard<std::mutex> lk(ctx().prof_mu);
    collect();
    ctx().prof = on != 0;
    return DMV_OK;
}

int profile_reset()
{
    DMV_TRY(ensure_device());
    std::lock_guard<std::mutex> lk(ctx().prof_mu);
    collect();
    Context &C = ctx();
    for (int i = 0; i < DMV_KC_COUNT; i++) {
        C.launches[i] = 0;
        C.ms[i] = 0;
        C.bytes[i] = 0;
    }
    return DMV_OK;
}

int profile_get(int kc, int64_t *launches, double *ms, double *bytes)
{
    DMV_CHECK_ARG(kc >= 0 && kc < DMV_KC_COUNT, "unknown kernel class %d", kc);
    DMV_TRY(ensure_device());
    std::lock_guard<std::mutex> lk(ctx().prof_mu);
    collect();
    Context &C = ctx();
    if (launches)
        *launches = C.launches[kc];
    if (ms)
        *ms = C.ms[kc];
    if (bytes)
        *bytes = C.bytes[kc];
    return DMV_OK;
}

void tune_set(const char *name, int64_t v)
{
    Context &C = ctx();
    std::lock_guard<std::mutex> lk(C.tune_mu);
    if (v < 0)
        C.tune.erase(name);  // unset: the default (or DMV_<NAME>) applies again
    else
        C.tune[name] = v;
    C.tune_epoch++;
}

uint64_t tune_epoch() { return ctx().tune_epoch.load(); }

}  // namespace dmv
