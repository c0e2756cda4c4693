// Fused tile passes, complex64.
#include "k_tile_impl.hpp"

namespace dmv {

int launch_tile_c64(int rlog, int tlog, bool nt, bool half, void *state, int n, StateLayout L, const DevPass *passes,
                    const DevPhase *phases, const DevOp *ops, uint32_t pass_idx)
{
    if (half) {
        if (tlog == 8 && rlog == 5) return run_tile<float, 5, 8, true>(nt, state, n, L, passes, phases, ops, pass_idx);
        return kNotHandled;
    }
    if (tlog == 8 && rlog == 5) return run_tile<float, 5, 8, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 8 && rlog == 4) return run_tile<float, 4, 8, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 8 && rlog == 3) return run_tile<float, 3, 8, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 6 && rlog == 4) return run_tile<float, 4, 6, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 6 && rlog == 3) return run_tile<float, 3, 6, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    return kNotHandled;
}

}  // namespace dmv
