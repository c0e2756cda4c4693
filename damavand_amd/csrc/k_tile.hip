// Fused tile kernels: one HBM pass applies a run of operations to every tile
// of 2^k amplitudes.  A workgroup of 2^TLOG threads holds its tile in
// registers, 2^RLOG amplitudes per thread; operations on register bits run in
// VGPRs, and the tile moves through LDS only when the set of register bits
// changes between phases.  The low qubits always sit in lanes at load and
// store time, so each group of 8 (complex128) or 16 (complex64) lanes reads
// or writes one full 128-byte line.
//
// This unit dispatches a pass to the shape-specific units (k_tile_c128*.hip,
// k_tile_c64.hip).
#include "k_tile_impl.hpp"

namespace dmv {

bool tile_half_exists(int dtype, int rlog, int tlog)
{
    return tlog == 8 && rlog == (dtype == DMV_C64 ? 5 : 4);
}

int launch_tile(int dtype, int rlog, int tlog, bool nt, bool half, void *state, int n, StateLayout L,
                const DevPass *passes, const DevPhase *phases, const DevOp *ops, uint32_t pass_idx)
{
    DMV_CHECK_ARG(n >= rlog + tlog && n - (rlog + tlog) <= 31, "tile shape does not fit %d qubits", n);
    DMV_CHECK_ARG(!half || tile_half_exists(dtype, rlog, tlog), "no half-LDS tile kernel for rlog %d tlog %d", rlog,
                  tlog);
    int rc = kNotHandled;
    if (dtype == DMV_C128) {
        rc = launch_tile_c128_main(rlog, tlog, nt, half, state, n, L, passes, phases, ops, pass_idx);
        if (rc == kNotHandled && !half)
            rc = launch_tile_c128_other(rlog, tlog, nt, state, n, L, passes, phases, ops, pass_idx);
    } else {
        rc = launch_tile_c64(rlog, tlog, nt, half, state, n, L, passes, phases, ops, pass_idx);
    }
    if (rc != kNotHandled)
        return rc;
    set_error("no tile kernel for dtype %d rlog %d tlog %d%s", dtype, rlog, tlog, half ? " (half LDS)" : "");
    return DMV_ERR_ARG;
}

}  // namespace dmv
