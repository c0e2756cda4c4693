// Kernel templates of the fused tile passes, shared by the translation
// units that instantiate them (one per dtype keeps each compile short).
#pragma once
#include "kernels.hpp"
#include "tile_ops.hpp"

namespace dmv {

// HALF: half a tile of LDS (split transposes), so that three workgroups of
// the 64 KiB tiles share a CU instead of two.
template <class R, int RLOG, int TLOG, bool NT, bool HALF>
__global__ void __launch_bounds__(1 << TLOG)
k_tile(typename Cplx<R>::T *__restrict__ state, const DevPass *__restrict__ passes,
       const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops, uint32_t pass_idx, uint32_t xcd_log2,
       StateLayout L)
{
    using V = typename Cplx<R>::T;
    __shared__ V lds[1 << (RLOG + TLOG - (HALF ? 1 : 0))];
    // Workgroups are dealt to the XCDs round-robin.  With xcd_log2 = x > 0
    // (grid a multiple of 2^x), XCD j takes the j-th contiguous 2^-x share of
    // the tiles, so the tiles one XCD runs at a time are neighbours in memory.
    const uint32_t b = blockIdx.x;
    const uint64_t tile = xcd_log2 ? (uint64_t)(b & ((1u << xcd_log2) - 1)) * (gridDim.x >> xcd_log2) + (b >> xcd_log2)
                                   : (uint64_t)b;
    tile_body<R, RLOG, TLOG, NT, HALF>(state, passes + pass_idx, phases, ops, tile, lds, L);
}

// ------------------------------------------------------------- launchers --
template <class R, int RLOG, int TLOG, bool HALF>
static int run_tile(bool nt, void *state, int n, StateLayout L, const DevPass *passes, const DevPhase *phases, const DevOp *ops,
                    uint32_t pass_idx)
{
    using V = typename Cplx<R>::T;
    const uint64_t tiles = uint64_t(1) << (n - (RLOG + TLOG));
    const int64_t xl = tune_value("tile_xcd_log2", 3);  // 8 XCDs per MI355X
    const uint32_t xcd_log2 = (xl > 0 && xl <= 4 && tiles >= (uint64_t(1) << xl)) ? (uint32_t)xl : 0u;
    if (nt)
        hipLaunchKernelGGL((k_tile<R, RLOG, TLOG, true, HALF>), dim3((uint32_t)tiles), dim3(1 << TLOG), 0, stream(),
                           (V *)state, passes, phases, ops, pass_idx, xcd_log2, L);
    else
        hipLaunchKernelGGL((k_tile<R, RLOG, TLOG, false, HALF>), dim3((uint32_t)tiles), dim3(1 << TLOG), 0,
                           stream(), (V *)state, passes, phases, ops, pass_idx, xcd_log2, L);
    return check_launch("tile pass");
}

// Shape-specific launchers (one translation unit each); they return
// kNotHandled when the shape (or its half-LDS variant) is not theirs.
constexpr int kNotHandled = -1;
int launch_tile_c128_main(int rlog, int tlog, bool nt, bool half, void *state, int n, StateLayout L,
                          const DevPass *passes, const DevPhase *phases, const DevOp *ops, uint32_t pass_idx);
int launch_tile_c128_other(int rlog, int tlog, bool nt, void *state, int n, StateLayout L, const DevPass *passes,
                           const DevPhase *phases, const DevOp *ops, uint32_t pass_idx);
int launch_tile_c128_wave(int rlog, int tlog, bool nt, void *state, int n, StateLayout L, const DevPass *passes,
                          const DevPhase *phases, const DevOp *ops, uint32_t pass_idx);
int launch_tile_c64(int rlog, int tlog, bool nt, bool half, void *state, int n, StateLayout L, const DevPass *passes,
                    const DevPhase *phases, const DevOp *ops, uint32_t pass_idx);

}  // namespace dmv
