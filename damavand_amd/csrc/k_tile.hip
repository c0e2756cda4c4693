// Fused tile kernels: one HBM pass applies a run of operations to every tile
// of 2^k amplitudes.  A workgroup of 2^TLOG threads holds its tile in
// registers, 2^RLOG amplitudes per thread; operations on register bits run in
// VGPRs, and the tile moves through LDS only when the set of register bits
// changes between phases.  The low qubits always sit in lanes at load and
// store time, so each group of 8 (complex128) or 16 (complex64) lanes reads
// or writes one full 128-byte line.
//
// Also here: the batched persistent kernel, in which each workgroup takes
// whole circuits from a queue and runs forward, lambda = H psi and the
// reverse sweep with the same tile bodies, serially over its own scratch
// states.
#include "kernels.hpp"
#include "tile_ops.hpp"

namespace dmv {

template <class R, int RLOG, int TLOG, bool NT>
__global__ void __launch_bounds__(1 << TLOG)
k_tile(typename Cplx<R>::T *__restrict__ state, const DevPass *__restrict__ passes,
       const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops, uint32_t pass_idx, uint32_t xcd_log2,
       StateLayout L)
{
    using V = typename Cplx<R>::T;
    __shared__ V lds[1 << (RLOG + TLOG)];
    // Workgroups are dealt to the XCDs round-robin.  With xcd_log2 = x > 0
    // (grid a multiple of 2^x), XCD j takes the j-th contiguous 2^-x share of
    // the tiles, so the tiles one XCD runs at a time are neighbours in memory.
    const uint32_t b = blockIdx.x;
    const uint64_t tile = xcd_log2 ? (uint64_t)(b & ((1u << xcd_log2) - 1)) * (gridDim.x >> xcd_log2) + (b >> xcd_log2)
                                   : (uint64_t)b;
    tile_body<R, RLOG, TLOG, NT>(state, passes + pass_idx, phases, ops, tile, lds, L);
}

// Persistent form: gridDim.x workgroups stride over the tiles with the next
// tile's loads in flight during the current tile's compute.
template <class R, int RLOG, int TLOG, bool NT>
__global__ void __launch_bounds__(1 << TLOG)
k_tile_pipe(typename Cplx<R>::T *__restrict__ state, const DevPass *__restrict__ passes,
            const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops, uint32_t pass_idx, uint64_t n_tiles,
            StateLayout L)
{
    using V = typename Cplx<R>::T;
    __shared__ V lds[1 << (RLOG + TLOG)];
    tile_stream<R, RLOG, TLOG, NT>(state, passes + pass_idx, phases, ops, blockIdx.x, gridDim.x, n_tiles, lds, L);
}

// partials[slot][tile][wave]: one value per wave of every tile.
struct PartialSink {
    static constexpr bool kPerWave = true;
    double *partials;
    uint64_t per_slot, base;  // per_slot = tiles x waves, base = tile x waves
    __device__ void operator()(int slot, double v, uint32_t wave) const
    {
        partials[(uint64_t)slot * per_slot + base + wave] = v;
    }
};

template <class R, int RLOG, int TLOG>
__global__ void __launch_bounds__(1 << TLOG)
k_tile_adjoint(typename Cplx<R>::T *__restrict__ psi, typename Cplx<R>::T *__restrict__ lam,
               const DevPass *__restrict__ passes, const DevPhase *__restrict__ phases,
               const DevOp *__restrict__ ops, uint32_t pass_idx, double *__restrict__ partials, StateLayout L)
{
    using V = typename Cplx<R>::T;
    __shared__ V lds[1 << (RLOG + TLOG)];
    __shared__ double red[(1 << TLOG) / 64];
    constexpr uint32_t W = (1u << TLOG) / 64;
    PartialSink sink{partials, (uint64_t)gridDim.x * W, (uint64_t)blockIdx.x * W};
    tile_body_adjoint<R, RLOG, TLOG>(psi, lam, passes + pass_idx, phases, ops, blockIdx.x, lds, red, sink, L);
}

// ------------------------------------------------------------- launchers --
template <class R, int RLOG, int TLOG>
static int run_tile(bool nt, void *state, int n, const DevPass *passes, const DevPhase *phases, const DevOp *ops,
                    uint32_t pass_idx)
{
    using V = typename Cplx<R>::T;
    const uint64_t tiles = uint64_t(1) << (n - (RLOG + TLOG));
    const StateLayout L = state_layout(n, sizeof(R) == 4);
    const int64_t per_cu = tune_value("tile_wgs_per_cu", 2);
    const int64_t forced = tune_value("tile_grid", 0);  // explicit persistent grid (tests)
    const uint64_t grid64 = forced > 0 ? (uint64_t)forced : (uint64_t)device_cus() * (uint64_t)(per_cu > 0 ? per_cu : 1);
    if (tune_value("tile_pipeline", 0) && per_cu > 0 && tiles >= 2 * grid64) {
        const uint32_t grid = (uint32_t)grid64;
        if (nt)
            hipLaunchKernelGGL((k_tile_pipe<R, RLOG, TLOG, true>), dim3(grid), dim3(1 << TLOG), 0, stream(),
                               (V *)state, passes, phases, ops, pass_idx, tiles, L);
        else
            hipLaunchKernelGGL((k_tile_pipe<R, RLOG, TLOG, false>), dim3(grid), dim3(1 << TLOG), 0, stream(),
                               (V *)state, passes, phases, ops, pass_idx, tiles, L);
        return check_launch("pipelined tile pass");
    }
    const int64_t xl = tune_value("tile_xcd_log2", 3);  // 8 XCDs per MI355X
    const uint32_t xcd_log2 = (xl > 0 && xl <= 4 && tiles >= (uint64_t(1) << xl)) ? (uint32_t)xl : 0u;
    if (nt)
        hipLaunchKernelGGL((k_tile<R, RLOG, TLOG, true>), dim3((uint32_t)tiles), dim3(1 << TLOG), 0, stream(),
                           (V *)state, passes, phases, ops, pass_idx, xcd_log2, L);
    else
        hipLaunchKernelGGL((k_tile<R, RLOG, TLOG, false>), dim3((uint32_t)tiles), dim3(1 << TLOG), 0, stream(),
                           (V *)state, passes, phases, ops, pass_idx, xcd_log2, L);
    return check_launch("tile pass");
}

int launch_tile(int dtype, int rlog, int tlog, bool nt, void *state, int n, const DevPass *passes,
                const DevPhase *phases, const DevOp *ops, uint32_t pass_idx)
{
    DMV_CHECK_ARG(n >= rlog + tlog && n - (rlog + tlog) <= 31, "tile shape does not fit %d qubits", n);
    if (dtype == DMV_C128) {
        if (tlog == 9 && rlog == 4) return run_tile<double, 4, 9>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 9 && rlog == 3) return run_tile<double, 3, 9>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 8 && rlog == 4) return run_tile<double, 4, 8>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 8 && rlog == 3) return run_tile<double, 3, 8>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 8 && rlog == 2) return run_tile<double, 2, 8>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 6 && rlog == 5) return run_tile<double, 5, 6>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 6 && rlog == 4) return run_tile<double, 4, 6>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 6 && rlog == 3) return run_tile<double, 3, 6>(nt, state, n, passes, phases, ops, pass_idx);
    } else {
        if (tlog == 8 && rlog == 5) return run_tile<float, 5, 8>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 8 && rlog == 4) return run_tile<float, 4, 8>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 8 && rlog == 3) return run_tile<float, 3, 8>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 6 && rlog == 4) return run_tile<float, 4, 6>(nt, state, n, passes, phases, ops, pass_idx);
        if (tlog == 6 && rlog == 3) return run_tile<float, 3, 6>(nt, state, n, passes, phases, ops, pass_idx);
    }
    set_error("no tile kernel for dtype %d rlog %d tlog %d", dtype, rlog, tlog);
    return DMV_ERR_ARG;
}

template <class R, int RLOG, int TLOG>
static int run_tile_adj(void *psi, void *lam, int n, const DevPass *passes, const DevPhase *phases,
                        const DevOp *ops, uint32_t pass_idx, double *partials)
{
    using V = typename Cplx<R>::T;
    const uint64_t tiles = uint64_t(1) << (n - (RLOG + TLOG));
    hipLaunchKernelGGL((k_tile_adjoint<R, RLOG, TLOG>), dim3((uint32_t)tiles), dim3(1 << TLOG), 0, stream(),
                       (V *)psi, (V *)lam, passes, phases, ops, pass_idx, partials,
                       state_layout(n, sizeof(R) == 4));
    return check_launch("adjoint tile pass");
}

int launch_tile_adjoint(int dtype, int rlog, int tlog, void *psi, void *lam, int n, const DevPass *passes,
                        const DevPhase *phases, const DevOp *ops, uint32_t pass_idx, double *partials)
{
    DMV_CHECK_ARG(n >= rlog + tlog && n - (rlog + tlog) <= 31, "tile shape does not fit %d qubits", n);
    if (dtype == DMV_C128) {
        if (tlog == 8 && rlog == 4) return run_tile_adj<double, 4, 8>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 8 && rlog == 3) return run_tile_adj<double, 3, 8>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 8 && rlog == 2) return run_tile_adj<double, 2, 8>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 6 && rlog == 3) return run_tile_adj<double, 3, 6>(psi, lam, n, passes, phases, ops, pass_idx, partials);
    } else {
        if (tlog == 8 && rlog == 5) return run_tile_adj<float, 5, 8>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 8 && rlog == 4) return run_tile_adj<float, 4, 8>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 8 && rlog == 3) return run_tile_adj<float, 3, 8>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 6 && rlog == 3) return run_tile_adj<float, 3, 6>(psi, lam, n, passes, phases, ops, pass_idx, partials);
        if (tlog == 6 && rlog == 4) return run_tile_adj<float, 4, 6>(psi, lam, n, passes, phases, ops, pass_idx, partials);
    }
    set_error("no adjoint tile kernel for dtype %d rlog %d tlog %d", dtype, rlog, tlog);
    return DMV_ERR_ARG;
}

// ---------------------------------------------------------------- batched --
struct GradSink {
    static constexpr bool kPerWave = false;
    double *grad;
    bool first;  // first tile of the pass: store, later tiles: accumulate
    __device__ void operator()(int slot, double v, uint32_t) const
    {
        if (first)
            grad[slot] = v;
        else
            grad[slot] += v;
    }
};

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

template <class R, int RLOG, int TLOG>
__global__ void __launch_bounds__(1 << TLOG)
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
                tile_body<R, RLOG, TLOG, false>(psi, passes + p, phases, ops, t, lds, kPlainLayout);
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
                tile_body_adjoint<R, RLOG, TLOG>(psi, lam, passes + p, phases, ops, t, lds, red, GradSink{g, t == 0},
                                                 kPlainLayout);
            __syncthreads();
        }
    }
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

template <class R, int RLOG, int TLOG>
static int run_batch(const BatchLaunch &L)
{
    using V = typename Cplx<R>::T;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream(), L.counter);
    DMV_TRY(check_launch("batch queue reset"));
    hipLaunchKernelGGL((k_batch<R, RLOG, TLOG>), dim3(L.grid), dim3(1 << TLOG), 0, stream(), L.circ, L.order, L.count,
                       L.counter, L.passes, L.phases, L.ops, L.sops, L.coef, (V *)L.scratch, L.slot_amps, L.energy,
                       L.grad);
    return check_launch("batch kernel");
}

int batch_tile_shape(int dtype, int *rlog, int *tlog)
{
    *tlog = 8;
    *rlog = dtype == DMV_C64 ? 5 : 4;
    return DMV_OK;
}

int launch_batch(int dtype, const BatchLaunch &L)
{
    DMV_CHECK_ARG(L.grid >= 1 && L.count >= 1, "empty batch");
    if (dtype == DMV_C64)
        return run_batch<float, 5, 8>(L);
    return run_batch<double, 4, 8>(L);
}

}  // namespace dmv
