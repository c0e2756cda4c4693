// Reverse-sweep tile passes of the adjoint gradient: psi and lambda move
// together, G^dagger is applied to both, and every parametrised gate's
// Im<lambda|P|psi> is reduced per wave into partials.
#include "kernels.hpp"
#include "tile_ops.hpp"

namespace dmv {

// partials[ip][tile][wave] for the IP ops of this pass (ip = the op's index
// among them): one value per wave of every tile, reduced by the host's
// per-pass slot sums, so the buffer holds one pass's IP ops only.
struct PartialSink {
    static constexpr bool kPerWave = true;
    double *partials;
    uint64_t per_ip, base;  // per_ip = tiles x waves, base = tile x waves
    __device__ void operator()(const DevOp *op, double v, uint32_t wave) const
    {
        partials[(uint64_t)op->ip_pos * per_ip + base + wave] = v;
    }
};

template <class R, int RLOG, int TLOG, bool NT>
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
    tile_body_adjoint<R, RLOG, TLOG, NT>(psi, lam, passes + pass_idx, phases, ops, blockIdx.x, lds, red, sink, L);
}

template <class R, int RLOG, int TLOG, bool NT = false>
static int run_tile_adj(void *psi, void *lam, int n, StateLayout L, const DevPass *passes, const DevPhase *phases,
                        const DevOp *ops, uint32_t pass_idx, double *partials)
{
    using V = typename Cplx<R>::T;
    const uint64_t tiles = uint64_t(1) << (n - (RLOG + TLOG));
    hipLaunchKernelGGL((k_tile_adjoint<R, RLOG, TLOG, NT>), dim3((uint32_t)tiles), dim3(1 << TLOG), 0, stream(),
                       (V *)psi, (V *)lam, passes, phases, ops, pass_idx, partials, L);
    return check_launch("adjoint tile pass");
}

int launch_tile_adjoint(int dtype, int rlog, int tlog, bool nt, void *psi, void *lam, int n, StateLayout L,
                        const DevPass *passes, const DevPhase *phases, const DevOp *ops, uint32_t pass_idx,
                        double *partials)
{
    DMV_CHECK_ARG(n >= rlog + tlog && n - (rlog + tlog) <= 31, "tile shape does not fit %d qubits", n);
    // psi and lambda are both held in registers: at most 8 (complex128)
    // or 16 (complex64) amplitudes per thread and state.  The shapes of
    // large states also come with streaming loads and stores; the others
    // (small states, or a tunable's choice) take the hint as a no-op.
    if (dtype == DMV_C128) {
        if (tlog == 8 && rlog == 3)
            return nt ? run_tile_adj<double, 3, 8, true>(psi, lam, n, L, passes, phases, ops, pass_idx, partials)
                      : run_tile_adj<double, 3, 8>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
        if (tlog == 8 && rlog == 2) return run_tile_adj<double, 2, 8>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
        if (tlog == 6 && rlog == 3) return run_tile_adj<double, 3, 6>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
    } else {
        if (tlog == 8 && rlog == 4)
            return nt ? run_tile_adj<float, 4, 8, true>(psi, lam, n, L, passes, phases, ops, pass_idx, partials)
                      : run_tile_adj<float, 4, 8>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
        if (tlog == 8 && rlog == 3)
            return nt ? run_tile_adj<float, 3, 8, true>(psi, lam, n, L, passes, phases, ops, pass_idx, partials)
                      : run_tile_adj<float, 3, 8>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
        if (tlog == 6 && rlog == 3) return run_tile_adj<float, 3, 6>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
        if (tlog == 6 && rlog == 4) return run_tile_adj<float, 4, 6>(psi, lam, n, L, passes, phases, ops, pass_idx, partials);
    }
    set_error("no adjoint tile kernel for dtype %d rlog %d tlog %d", dtype, rlog, tlog);
    return DMV_ERR_ARG;
}

}  // namespace dmv
