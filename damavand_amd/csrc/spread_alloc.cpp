// Physically spread state buffers.
//
// On MI355X a streaming pass over a state held in one physically contiguous
// range runs at about 5.4 TB/s for every size up to 64 GiB, while the same
// pass over a state whose eighths lie 16 GiB apart in physical memory runs at
// about 6.2 TB/s.  Measured with tile-shaped in-place copies of a 16 GiB
// state: low-qubit tiles 5.4 -> 6.2 TB/s, tiles over qubits 21-29 5.0 -> 5.6
// TB/s; a 128 GiB state, whose eighths are 16 GiB apart anyway, runs at
// 6.2 TB/s from a plain allocation.  Eighths 8 GiB apart do not help, and
// neither does spacing them 16 GiB apart in virtual addresses only.
//
// The buffer is therefore built from eight physical pieces.  At allocation
// time each piece is followed by a spacer that brings the stride to 16 GiB
// (allocating piece, spacer, piece, ... in that order places the pieces at
// the stride: measured), the spacers are released at once, and the pieces are
// mapped back to back into one reserved address range, so that kernels and
// copies see one ordinary buffer.  The cost is 128 GiB of free device memory
// during the allocation only.  Without it, or for buffers under 2 GiB or of
// 128 GiB and more, the caller makes a plain allocation.
#include "runtime.hpp"

#include "device.hpp"

namespace dmv {

namespace {
constexpr uint64_t kGiB = uint64_t(1) << 30;
constexpr uint64_t kSpreadStride = 16 * kGiB;
}  // namespace

int spread_alloc(uint64_t bytes, SpreadBuffer *out)
{
    *out = SpreadBuffer();
    const uint64_t min_bytes = (uint64_t)tune_value("spread_min_bytes", (int64_t)(2 * kGiB));
    if (bytes < min_bytes || bytes % kSpreadPieces)
        return DMV_ERR_MEMORY;
    const uint64_t piece = bytes / kSpreadPieces;
    if (piece >= kSpreadStride)
        return DMV_ERR_MEMORY;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return DMV_ERR_MEMORY;
    }
    hipMemAllocationProp prop = {};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = dev;
    size_t gran = 0, free_b = 0, total_b = 0;
    if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended) != hipSuccess ||
        gran == 0 || piece % gran || (kSpreadStride - piece) % gran ||
        hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < kSpreadPieces * kSpreadStride + kGiB) {
        (void)hipGetLastError();
        return DMV_ERR_MEMORY;
    }
    hipMemGenericAllocationHandle_t spacer[kSpreadPieces] = {};
    int made = 0, spacers = 0, mapped = 0;
    bool ok = true;
    for (int c = 0; c < kSpreadPieces && ok; c++) {
        ok = hipMemCreate(&out->h[c], piece, &prop, 0) == hipSuccess;
        if (ok)
            made++;
        if (ok)
            ok = hipMemCreate(&spacer[c], kSpreadStride - piece, &prop, 0) == hipSuccess;
        if (ok)
            spacers++;
    }
    for (int c = 0; c < spacers; c++)
        (void)hipMemRelease(spacer[c]);
    void *va = nullptr;
    if (ok)
        ok = hipMemAddressReserve(&va, bytes, gran, nullptr, 0) == hipSuccess;
    hipMemAccessDesc acc = {};
    acc.location.type = hipMemLocationTypeDevice;
    acc.location.id = dev;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    for (int c = 0; c < kSpreadPieces && ok; c++) {
        char *at = (char *)va + (uint64_t)c * piece;
        ok = hipMemMap(at, piece, 0, out->h[c], 0) == hipSuccess;
        if (ok)
            mapped++;
        if (ok)
            ok = hipMemSetAccess(at, piece, &acc, 1) == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        for (int c = 0; c < mapped; c++)
            (void)hipMemUnmap((char *)va + (uint64_t)c * piece, piece);
        if (va)
            (void)hipMemAddressFree(va, bytes);
        for (int c = 0; c < made; c++)
            (void)hipMemRelease(out->h[c]);
        (void)hipGetLastError();
        *out = SpreadBuffer();
        return DMV_ERR_MEMORY;
    }
    out->va = va;
    out->bytes = bytes;
    out->piece = piece;
    return DMV_OK;
}

void spread_free(SpreadBuffer *b)
{
    if (!b->va)
        return;
    (void)hipStreamSynchronize(stream());
    for (int c = 0; c < kSpreadPieces; c++)
        (void)hipMemUnmap((char *)b->va + (uint64_t)c * b->piece, b->piece);
    (void)hipMemAddressFree(b->va, b->bytes);
    for (int c = 0; c < kSpreadPieces; c++)
        (void)hipMemRelease(b->h[c]);
    *b = SpreadBuffer();
}

}  // namespace dmv
