// Whole-state kernels for small registers: one workgroup loads the full state
// (<= 64 KiB) into LDS, applies every op of the list in order, and writes it
// back, so a forward of a few-qubit circuit is a single launch.
#include "kernels.hpp"
#include "tile_ops.hpp"

namespace dmv {

constexpr int kSmallThreads = 256;

template <class R>
__global__ void __launch_bounds__(kSmallThreads)
k_small(typename Cplx<R>::T *__restrict__ state, uint32_t n, const DevOp *__restrict__ ops, uint32_t n_ops)
{
    using V = typename Cplx<R>::T;
    constexpr int CAP = 65536 / sizeof(V);
    __shared__ V s[CAP];
    const uint32_t n_amp = 1u << n;
    for (uint32_t i = threadIdx.x; i < n_amp; i += blockDim.x)
        s[i] = state[i];
    __syncthreads();
    for (uint32_t o = 0; o < n_ops; o++) {
        small_apply<V, R>(s, n, ops + o);
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < n_amp; i += blockDim.x)
        state[i] = s[i];
}

// Reverse sweep with both states in LDS (each <= 32 KiB).
template <class R>
__global__ void __launch_bounds__(kSmallThreads)
k_small_adjoint(typename Cplx<R>::T *__restrict__ psi, typename Cplx<R>::T *__restrict__ lam, uint32_t n,
                const DevOp *__restrict__ ops, uint32_t n_ops, double *__restrict__ partials)
{
    using V = typename Cplx<R>::T;
    constexpr int CAP = 32768 / sizeof(V);
    __shared__ V sp[CAP];
    __shared__ V sl[CAP];
    __shared__ double red[kSmallThreads / 64];
    const uint32_t n_amp = 1u << n;
    for (uint32_t i = threadIdx.x; i < n_amp; i += blockDim.x) {
        sp[i] = psi[i];
        sl[i] = lam[i];
    }
    __syncthreads();
    for (uint32_t o = 0; o < n_ops; o++) {
        const DevOp *op = ops + o;
        if (op_is_ip(op->kind)) {
            const double s = block_sum<kSmallThreads>(small_ip(sp, sl, n, op), red);
            if (threadIdx.x == 0)
                partials[op->slot] = s;
        } else {
            small_apply<V, R>(sp, n, op);
            small_apply<V, R>(sl, n, op);
        }
        __syncthreads();
    }
    for (uint32_t i = threadIdx.x; i < n_amp; i += blockDim.x) {
        psi[i] = sp[i];
        lam[i] = sl[i];
    }
}

int launch_small(int dtype, void *state, int n, const DevOp *ops, uint32_t n_ops)
{
    const int cap = dtype == DMV_C64 ? kSmallMaxQubitsC64 : kSmallMaxQubitsC128;
    DMV_CHECK_ARG(n >= 1 && n <= cap, "whole-state kernel holds at most %d qubits", cap);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_small<float>, dim3(1), dim3(kSmallThreads), 0, stream(), (c64 *)state, (uint32_t)n, ops,
                           n_ops);
    else
        hipLaunchKernelGGL(k_small<double>, dim3(1), dim3(kSmallThreads), 0, stream(), (c128 *)state, (uint32_t)n,
                           ops, n_ops);
    return check_launch("whole-state kernel");
}

int launch_small_adjoint(int dtype, void *psi, void *lam, int n, const DevOp *ops, uint32_t n_ops, double *partials)
{
    const int cap = (dtype == DMV_C64 ? kSmallMaxQubitsC64 : kSmallMaxQubitsC128) - 1;
    DMV_CHECK_ARG(n >= 1 && n <= cap, "whole-state adjoint kernel holds at most %d qubits", cap);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_small_adjoint<float>, dim3(1), dim3(kSmallThreads), 0, stream(), (c64 *)psi, (c64 *)lam,
                           (uint32_t)n, ops, n_ops, partials);
    else
        hipLaunchKernelGGL(k_small_adjoint<double>, dim3(1), dim3(kSmallThreads), 0, stream(), (c128 *)psi,
                           (c128 *)lam, (uint32_t)n, ops, n_ops, partials);
    return check_launch("whole-state adjoint kernel");
}

}  // namespace dmv
