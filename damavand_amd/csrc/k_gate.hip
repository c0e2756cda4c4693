// Single-gate kernels: one pass over exactly the amplitudes a gate changes.
// Thread work items are pair (or amplitude) indices; the index of the first
// amplitude is the item index with a zero bit inserted at every target and
// control position and the control bits then set, so consecutive lanes walk
// consecutive addresses in each of the two streams (a[i0], a[i0 + 2^t]).
#include "device.hpp"
#include "kernels.hpp"

namespace dmv {

constexpr int kGateThreads = 256;

enum GateMode { GM_GEN = 0, GM_SWAP = 1, GM_DIAG = 2, GM_PHASE = 3 };

struct Positions {
    uint32_t pos[8];   // ascending bit positions to insert zeros at
    uint32_t count;
};

__device__ __forceinline__ uint64_t expand(uint64_t p, const Positions &P)
{
    for (uint32_t j = 0; j < P.count; j++)
        p = insert_zero_bit(p, P.pos[j]);
    return p;
}

template <class R, int PPT, int MODE, bool NT>
__global__ void __launch_bounds__(kGateThreads)
k_gate(typename Cplx<R>::T *__restrict__ state, uint64_t n_items, uint32_t t, Positions P, uint64_t ormask,
       Mat<R> M, StateLayout L)
{
    using V = typename Cplx<R>::T;
    const uint64_t base = (uint64_t)blockIdx.x * (kGateThreads * PPT) + threadIdx.x;
    uint64_t idx[PPT], e[PPT];
    bool ok[PPT];
#pragma unroll
    for (int u = 0; u < PPT; u++) {
        const uint64_t p = base + (uint64_t)u * kGateThreads;
        ok[u] = p < n_items;
        idx[u] = expand(p, P) | ormask;
        e[u] = elem(idx[u], L);
    }
    // idx has bit t clear, so the partner's element is e ^ elem(2^t)
    const uint64_t tb = elem(uint64_t(1) << t, L);
    if constexpr (MODE == GM_GEN || MODE == GM_SWAP) {
        V a[PPT], b[PPT];
#pragma unroll
        for (int u = 0; u < PPT; u++)
            if (ok[u]) {
                a[u] = ld<NT>(state + e[u]);
                b[u] = ld<NT>(state + (e[u] ^ tb));
            }
#pragma unroll
        for (int u = 0; u < PPT; u++)
            if (ok[u]) {
                if constexpr (MODE == GM_GEN) {
                    apply2(a[u], b[u], M);
                    st<NT>(state + e[u], a[u]);
                    st<NT>(state + (e[u] ^ tb), b[u]);
                } else {
                    st<NT>(state + e[u], b[u]);
                    st<NT>(state + (e[u] ^ tb), a[u]);
                }
            }
    } else {
        V a[PPT];
#pragma unroll
        for (int u = 0; u < PPT; u++)
            if (ok[u])
                a[u] = ld<NT>(state + e[u]);
#pragma unroll
        for (int u = 0; u < PPT; u++)
            if (ok[u]) {
                if constexpr (MODE == GM_DIAG) {
                    const bool one = (idx[u] >> t) & 1u;
                    a[u] = cmul(a[u], one ? M.m11r : M.m00r, one ? M.m11i : M.m00i);
                } else {
                    a[u] = cmul(a[u], M.m11r, M.m11i);
                }
                st<NT>(state + e[u], a[u]);
            }
    }
}

template <class R, int PPT, int MODE>
static int run_gate(bool nt, void *state, uint64_t items, uint32_t t, const Positions &P, uint64_t ormask,
                    const Mat<R> &M, StateLayout L)
{
    using V = typename Cplx<R>::T;
    const uint64_t per_block = (uint64_t)kGateThreads * PPT;
    const uint64_t blocks = (items + per_block - 1) / per_block;
    DMV_CHECK_ARG(blocks > 0 && blocks <= 0xFFFFFFFFull, "gate grid out of range");
    if (nt)
        hipLaunchKernelGGL((k_gate<R, PPT, MODE, true>), dim3((uint32_t)blocks), dim3(kGateThreads), 0, stream(),
                           (V *)state, items, t, P, ormask, M, L);
    else
        hipLaunchKernelGGL((k_gate<R, PPT, MODE, false>), dim3((uint32_t)blocks), dim3(kGateThreads), 0,
                           stream(), (V *)state, items, t, P, ormask, M, L);
    return check_launch("gate kernel");
}

template <class R, int MODE>
static int run_gate_ppt(int ppt, bool nt, void *state, uint64_t items, uint32_t t, const Positions &P,
                        uint64_t ormask, const Mat<R> &M, StateLayout L)
{
    if (ppt >= 4)
        return run_gate<R, 4, MODE>(nt, state, items, t, P, ormask, M, L);
    if (ppt == 2)
        return run_gate<R, 2, MODE>(nt, state, items, t, P, ormask, M, L);
    return run_gate<R, 1, MODE>(nt, state, items, t, P, ormask, M, L);
}

template <class R>
static int gate_dispatch(int mode, int ppt, bool nt, void *state, uint64_t items, uint32_t t,
                         const Positions &P, uint64_t ormask, const Mat<R> &M, StateLayout L)
{
    switch (mode) {
    case GM_GEN: return run_gate_ppt<R, GM_GEN>(ppt, nt, state, items, t, P, ormask, M, L);
    case GM_SWAP: return run_gate_ppt<R, GM_SWAP>(ppt, nt, state, items, t, P, ormask, M, L);
    case GM_DIAG: return run_gate_ppt<R, GM_DIAG>(ppt, nt, state, items, t, P, ormask, M, L);
    default: return run_gate_ppt<R, GM_PHASE>(ppt, nt, state, items, t, P, ormask, M, L);
    }
}

int launch_gate(int dtype, void *state, int n, StateLayout L, const DevOp &op, int ppt, bool nt)
{
    DMV_CHECK_ARG(op.target < (uint32_t)n, "target %u out of range", op.target);
    DMV_CHECK_ARG((op.ctrl >> n) == 0 && !((op.ctrl >> op.target) & 1), "bad control mask");
    DMV_CHECK_ARG(popcount64(op.ctrl) <= 6, "at most 6 controls");
    const uint64_t tb = uint64_t(1) << op.target;
    const bool phase_only = op.kind == OP_DIAG && op.m[0] == 1.0 && op.m[1] == 0.0;
    int mode;
    uint64_t ins, ormask;
    if (op.kind == OP_GEN || op.kind == OP_SWAP) {
        mode = op.kind == OP_GEN ? GM_GEN : GM_SWAP;
        ins = op.ctrl | tb;
        ormask = op.ctrl;
    } else if (op.kind == OP_DIAG && phase_only) {
        mode = GM_PHASE;
        ins = op.ctrl | tb;
        ormask = op.ctrl | tb;
    } else if (op.kind == OP_DIAG) {
        mode = GM_DIAG;
        ins = op.ctrl;
        ormask = op.ctrl;
    } else {
        set_error("op kind %u has no single-gate kernel", op.kind);
        return DMV_ERR_ARG;
    }
    Positions P;
    P.count = 0;
    for (int q = 0; q < n; q++)
        if ((ins >> q) & 1)
            P.pos[P.count++] = (uint32_t)q;
    // one work item per changed pair (GEN/SWAP) or amplitude (DIAG/PHASE)
    const uint64_t items = (uint64_t(1) << n) >> P.count;
    if (dtype == DMV_C64)
        return gate_dispatch<float>(mode, ppt, nt, state, items, op.target, P, ormask, load_mat<float>(op.m), L);
    return gate_dispatch<double>(mode, ppt, nt, state, items, op.target, P, ormask, load_mat<double>(op.m), L);
}

}  // namespace dmv
