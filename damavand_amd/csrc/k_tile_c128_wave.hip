// Fused tile passes, complex128, one-wave (64-thread) shapes for small states.
#include "k_tile_impl.hpp"

namespace dmv {

int launch_tile_c128_wave(int rlog, int tlog, bool nt, void *state, int n, StateLayout L, const DevPass *passes, const DevPhase *phases,
                          const DevOp *ops, uint32_t pass_idx)
{
    if (tlog == 6 && rlog == 5) return run_tile<double, 5, 6, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 6 && rlog == 4) return run_tile<double, 4, 6, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 6 && rlog == 3) return run_tile<double, 3, 6, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    return kNotHandled;
}

}  // namespace dmv
