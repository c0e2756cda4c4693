// Fused tile passes, complex128, the default 16-register shapes.
#include "k_tile_impl.hpp"

namespace dmv {

int launch_tile_c128_main(int rlog, int tlog, bool nt, bool half, void *state, int n, StateLayout L,
                          const DevPass *passes, const DevPhase *phases, const DevOp *ops, uint32_t pass_idx)
{
    if (half) {
        if (tlog == 8 && rlog == 4) return run_tile<double, 4, 8, true>(nt, state, n, L, passes, phases, ops, pass_idx);
        return kNotHandled;
    }
    if (tlog == 9 && rlog == 4) return run_tile<double, 4, 9, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 8 && rlog == 4) return run_tile<double, 4, 8, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    return kNotHandled;
}

}  // namespace dmv
