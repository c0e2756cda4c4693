// The batched persistent kernel: each workgroup takes whole circuits from a
// queue and runs forward, lambda = H psi and the reverse sweep with the same
// tile bodies, serially over its own scratch states.
#include "kernels.hpp"
#include "tile_ops.hpp"

namespace dmv {

// ---------------------------------------------------------------- batched --
// Gradient partials of a reverse pass over a circuit's tiles: each wave
// keeps its own sum per IP op of the pass in LDS (acc[ip_pos][wave], first
// tile stores, later tiles add, in tile order), so no workgroup barrier is
// needed per output; after the pass thread 0 adds the waves' sums in wave
// order (deterministic) into the gradient.
struct WaveAccSink {
    static constexpr bool kPerWave = true;
    double *acc;
    uint32_t waves;
    bool first;
    __device__ void operator()(const DevOp *op, double v, uint32_t wave) const
    {
        double *a = acc + op->ip_pos * waves + wave;
        *a = first ? v : *a + v;
    }
};

// After a reverse pass: the gradient entries of its IP ops (thread 0).
__device__ __forceinline__ void flush_pass_grads(const DevPass *P, const DevPhase *__restrict__ phases,
                                                 const DevOp *__restrict__ ops, const double *acc, uint32_t waves,
                                                 double *g)
{
    for (uint32_t ph = P->phase_begin; ph < P->phase_end; ph++)
        for (uint32_t o = phases[ph].op_begin; o < phases[ph].op_end; o++) {
            const DevOp *op = ops + o;
            if (!op_is_ip(op->kind))
                continue;
            double s = 0.0;
            for (uint32_t w = 0; w < waves; w++)
                s += acc[op->ip_pos * waves + w];
            g[op->slot] = s;
        }
}

// h(x) of the chain observable sum_q a_q z_q + sum_{q<n-1} b_q z_q z_{q+1}.
__device__ __forceinline__ double chain_h(uint64_t x, uint32_t n, const double *a, const double *b)
{
    double h = 0.0;
    for (uint32_t q = 0; q < n; q++) {
        const uint32_t bq = (x >> q) & 1u;
        h += bq ? -a[q] : a[q];
        if (q + 1 < n)
            h += (bq ^ ((x >> (q + 1)) & 1u)) ? -b[q] : b[q];
    }
    return h;
}

// At most 168 VGPRs, so that three 256-thread workgroups share a CU.
// STREAM: streaming loads and stores of the tile path's scratch states.
template <class R, int RLOG, int TLOG, bool STREAM>
__global__ void __launch_bounds__(1 << TLOG) __attribute__((amdgpu_waves_per_eu(3)))
k_batch(const BatchCircuit *__restrict__ circ, const uint32_t *__restrict__ order, uint32_t count,
        uint32_t *__restrict__ counter, const DevPass *__restrict__ passes, const DevPhase *__restrict__ phases,
        const DevOp *__restrict__ ops, const DevOp *__restrict__ sops, const double *__restrict__ coef,
        typename Cplx<R>::T *__restrict__ scratch, uint64_t slot_amps, double *__restrict__ energy,
        double *__restrict__ grad)
{
    using V = typename Cplx<R>::T;
    constexpr int K = RLOG + TLOG, NT = 1 << TLOG;
    __shared__ V lds[1 << K];
    __shared__ double red[NT / 64];
    __shared__ double acc[kBatchMaxIp * (NT / 64)];
    __shared__ uint32_t s_idx;
    const uint32_t tid = threadIdx.x;
    V *psi = scratch + (uint64_t)blockIdx.x * 2 * slot_amps;
    V *lam = psi + slot_amps;
    for (;;) {
        if (tid == 0)
            s_idx = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t w = s_idx;
        __syncthreads();
        if (w >= count)
            break;
        const uint32_t ci = order[w];
        const BatchCircuit C = circ[ci];
        const uint32_t n = C.n;
        const uint64_t n_amp = uint64_t(1) << n;
        const double *a = coef + C.coef;
        const double *b = a + n;
        double *g = grad + C.grad;
        if (C.small) {
            // psi and lambda live in the two halves of LDS
            V *sp = lds, *sl = lds + (1 << (K - 1));
            for (uint32_t i = tid; i < n_amp; i += NT) {
                V z;
                z.x = i == 0 ? 1 : 0;
                z.y = 0;
                sp[i] = z;
            }
            __syncthreads();
            for (uint32_t o = C.fwd_begin; o < C.fwd_end; o++) {
                small_apply<V, R>(sp, n, sops + o);
                __syncthreads();
            }
            double e = 0.0;
            for (uint32_t i = tid; i < n_amp; i += NT) {
                const double h = chain_h(i, n, a, b);
                const V v = sp[i];
                e += h * norm2(v);
                V w2;
                w2.x = v.x * (R)h;
                w2.y = v.y * (R)h;
                sl[i] = w2;
            }
            e = block_sum<NT>(e, red);
            if (tid == 0)
                energy[ci] = e;
            __syncthreads();
            for (uint32_t o = C.adj_begin; o < C.adj_end; o++) {
                const DevOp *op = sops + o;
                if (op_is_ip(op->kind)) {
                    const double s = block_sum<NT>(small_ip(sp, sl, n, op), red);
                    if (tid == 0)
                        g[op->slot] = s;
                } else {
                    small_apply<V, R>(sp, n, op);
                    small_apply<V, R>(sl, n, op);
                }
                __syncthreads();
            }
            continue;
        }
        // tile path over the workgroup's scratch states
        for (uint64_t i = tid; i < n_amp; i += NT) {
            V z;
            z.x = i == 0 ? 1 : 0;
            z.y = 0;
            psi[i] = z;
        }
        __syncthreads();
        const uint64_t n_tiles = n_amp >> K;
        for (uint32_t p = C.fwd_begin; p < C.fwd_end; p++) {
            for (uint64_t t = 0; t < n_tiles; t++)
                tile_body<R, RLOG, TLOG, STREAM>(psi, passes + p, phases, ops, t, lds, kPlainLayout);
            __syncthreads();
        }
        double e = 0.0;
        for (uint64_t i = tid; i < n_amp; i += NT) {
            const double h = chain_h(i, n, a, b);
            const V v = psi[i];
            e += h * norm2(v);
            V w2;
            w2.x = v.x * (R)h;
            w2.y = v.y * (R)h;
            lam[i] = w2;
        }
        e = block_sum<NT>(e, red);
        if (tid == 0)
            energy[ci] = e;
        __syncthreads();
        for (uint32_t p = C.adj_begin; p < C.adj_end; p++) {
            for (uint64_t t = 0; t < n_tiles; t++)
                tile_body_adjoint<R, RLOG, TLOG, STREAM>(psi, lam, passes + p, phases, ops, t, lds, red,
                                                 WaveAccSink{acc, NT / 64, t == 0}, kPlainLayout);
            __syncthreads();
            if (tid == 0)
                flush_pass_grads(passes + p, phases, ops, acc, NT / 64, g);
            __syncthreads();  // the next pass's first tile overwrites acc
        }
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

template <class R, int RLOG, int TLOG, bool STREAM = false>
static int run_batch(const BatchLaunch &L)
{
    using V = typename Cplx<R>::T;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream(), L.counter);
    DMV_TRY(check_launch("batch queue reset"));
    hipLaunchKernelGGL((k_batch<R, RLOG, TLOG, STREAM>), dim3(L.grid), dim3(1 << TLOG), 0, stream(), L.circ, L.order, L.count,
                       L.counter, L.passes, L.phases, L.ops, L.sops, L.coef, (V *)L.scratch, L.slot_amps, L.energy,
                       L.grad);
    return check_launch("batch kernel");
}

// Tile shapes (log2 amplitudes per thread and state, log2 threads) of the
// batched kernel: 8 amplitudes per thread and state at most, since the
// reverse sweep holds psi and lambda in registers together.
#define DMV_BATCH_SHAPES(X) X(3, 8) X(2, 8) X(3, 7) X(3, 9)

bool batch_kernel_exists(int dtype, int rlog, int tlog)
{
    (void)dtype;
#define DMV_BATCH_HAS(R, T)                                                                        \
    if (rlog == R && tlog == T)                                                                    \
        return true;
    DMV_BATCH_SHAPES(DMV_BATCH_HAS)
#undef DMV_BATCH_HAS
    return false;
}

int launch_batch(int dtype, const BatchLaunch &L)
{
    DMV_CHECK_ARG(L.grid >= 1 && L.count >= 1, "empty batch");
    // the default shape also comes with streaming loads and stores; the
    // other shapes (tunables) take the hint as a no-op
    if (L.nt && L.rlog == 3 && L.tlog == 8)
        return dtype == DMV_C64 ? run_batch<float, 3, 8, true>(L) : run_batch<double, 3, 8, true>(L);
#define DMV_BATCH_RUN(R, T)                                                                        \
    if (L.rlog == R && L.tlog == T)                                                                \
        return dtype == DMV_C64 ? run_batch<float, R, T>(L) : run_batch<double, R, T>(L);
    DMV_BATCH_SHAPES(DMV_BATCH_RUN)
#undef DMV_BATCH_RUN
    set_error("no batched kernel for rlog %d tlog %d", L.rlog, L.tlog);
    return DMV_ERR_ARG;
}

}  // namespace dmv
