// Fused tile passes, complex128, the other 256/512/128-thread shapes.
#include "k_tile_impl.hpp"

namespace dmv {

int launch_tile_c128_other(int rlog, int tlog, bool nt, void *state, int n, StateLayout L, const DevPass *passes, const DevPhase *phases,
                           const DevOp *ops, uint32_t pass_idx)
{
    if (tlog == 9 && rlog == 3) return run_tile<double, 3, 9, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 8 && rlog == 3) return run_tile<double, 3, 8, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 7 && rlog == 4) return run_tile<double, 4, 7, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    if (tlog == 8 && rlog == 2) return run_tile<double, 2, 8, false>(nt, state, n, L, passes, phases, ops, pass_idx);
    return launch_tile_c128_wave(rlog, tlog, nt, state, n, L, passes, phases, ops, pass_idx);
}

}  // namespace dmv
