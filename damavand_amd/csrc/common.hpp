// Shared definitions of libdamavand-gpu: status codes, error reporting, the
// device-side plan descriptors consumed by the tile kernels, and small bit
// helpers used by both host and device code.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/damavand_gpu.h"

#if defined(__HIPCC__)
#define DMV_HD __host__ __device__
#else
#define DMV_HD
#endif

namespace dmv {

// ---------------------------------------------------------------- errors --
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
const char *last_error();

#define DMV_CHECK_ARG(cond, ...)                                                                   \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            ::dmv::set_error(__VA_ARGS__);                                                         \
            return DMV_ERR_ARG;                                                                    \
        }                                                                                          \
    } while (0)

#define DMV_TRY(expr)                                                                              \
    do {                                                                                           \
        int rc_ = (expr);                                                                          \
        if (rc_ != DMV_OK)                                                                         \
            return rc_;                                                                            \
    } while (0)

// ------------------------------------------------------------ constants --
constexpr int kMaxQubits = 33;      // 2^33 complex128 = 128 GiB, one MI355X
constexpr int kMaxRegs = 32;        // amplitudes per thread in a tile kernel
constexpr int kMaxThreadBits = 9;   // up to 512 threads per tile workgroup
constexpr int kMaxTileQubits = 13;  // 2^13 * 8 B (c64) or 2^12 * 16 B (c128) = 64 KiB of LDS
constexpr int kSmallMaxQubitsC128 = 12;  // whole state in one workgroup's LDS (64 KiB)
constexpr int kSmallMaxQubitsC64 = 13;

inline int amp_bytes(int dtype) { return dtype == DMV_C64 ? 8 : 16; }
// Low qubits a tile always keeps in lanes so that 8 consecutive lanes read one
// full 128-byte line: 3 for complex128 (8 x 16 B), 4 for complex64 (16 x 8 B).
inline int coalesce_bits(int dtype) { return dtype == DMV_C64 ? 4 : 3; }

// ------------------------------------------------------ operation kinds --
// A tile/gate operation.  GEN: 2x2 matrix on a register-resident target.
// DIAG: diag(m00, m11) on any bit.  SWAP: X on a register-resident target.
// IP ops (adjoint): Im<lambda|A|psi> for A = ax X + ay Y + az Z on the
// target, with (ax, ay, az) in m[0..2]; IPG on a register-resident target,
// IPZ (ax = ay = 0) on any bit.
enum OpKind : uint32_t {
    OP_GEN = 0,
    OP_DIAG = 1,
    OP_SWAP = 2,
    OP_IPG = 4,
    OP_IPZ = 6,
};

inline DMV_HD bool op_needs_register(uint32_t kind) { return kind == OP_GEN || kind == OP_SWAP || kind == OP_IPG; }
inline DMV_HD bool op_is_ip(uint32_t kind) { return kind == OP_IPG || kind == OP_IPZ; }

constexpr uint32_t kNoReg = 0xFFFFFFFFu;

// One operation as the kernels read it.  For single-gate and whole-state
// kernels `target` and `ctrl` are global qubit positions and nothing else is
// used.  In tile plans they are local tile bits (kept for checking) and the
// kernel uses the host-resolved fields:
//  * rmask: for pair ops (GEN, SWAP, IPX, IPY) the physical register partner
//    mask (register p pairs with p ^ rmask); for a diagonal or IPZ whose
//    target is a register bit, the set of physical registers whose target
//    bit is 1;
//  * fl: pair ops: the physical registers holding the target-bit-1 member;
//  * en: the physical registers whose register-bit controls hold;
//  * tbit: the thread bit of a diagonal/IPZ target that is not a register
//    bit (kNoReg otherwise);
//  * ctrl_thr: controls that are thread bits (mask over the thread id);
//  * plain: a GEN op acting on every register (en = all) whose first pair
//    members (registers with the partner mask's highest bit clear) all hold
//    target bit 0, so the matrix applies without per-pair selects; for a
//    DIAG op: the diagonal is exactly diag(1, -1) (Z, CZ), so the op only
//    flips the signs of the amplitudes whose target bit is 1; for an IP
//    op: the previous op of the phase is an IP op of the same kind, target
//    and masks, so its per-thread sums of X, Y and Z are reused.
struct alignas(16) DevOp {
    uint32_t kind;
    uint32_t target;
    uint32_t rmask;
    uint32_t tbit;
    int32_t slot;      // output slot of an IP op (parameter index), else -1
    uint32_t en;
    uint32_t fl;
    uint32_t ctrl_thr;
    uint64_t ctrl;     // every bit of ctrl must be 1 for the op to act
    uint32_t plain;
    uint32_t ip_pos;   // IP op of a tile plan: its index among the IP ops of its pass
    double m[8];       // re/im of m00, m01, m10, m11
};
static_assert(sizeof(DevOp) == 112, "DevOp layout");

// Register/thread assignment of a tile's k local bits: the local index of
// register r in thread t is lreg[r] | sum_j bit_j(t) << thr[j].  lswz[r] is
// the LDS byte offset of register r's part of the (linear) swizzle, so the
// LDS address of (t, r) is swz(thread part) ^ lswz[r].
struct alignas(16) DevLayout {
    uint32_t lreg[kMaxRegs];
    uint32_t lswz[kMaxRegs];
    uint32_t thr[kMaxThreadBits];
};

// LDS address swizzle of a local index: folds the higher G-bit groups into
// the low G bits.  A local bit p lands on swizzled bit p mod G, so lanes that
// differ in bits of distinct residues never share a 128-bit bank group.  It is
// linear over GF(2): swz(a ^ b) = swz(a) ^ swz(b).
template <int G> inline DMV_HD uint32_t swz_index(uint32_t li)
{
    uint32_t f = 0;
    for (int s = G; s < 16; s += G)
        f ^= li >> s;
    return li ^ (f & ((1u << G) - 1));
}

// LDS offsets of a transpose: register p is written at swz(thread part in the
// source layout) ^ src_lswz[p] and register r read back at swz(thread part in
// the destination layout) ^ dst_lswz[r].  In a pass with a split (DevPass),
// both sides hold the offsets with the split bit removed.
struct alignas(16) DevPhase {
    DevLayout lay;
    uint32_t src_lswz[kMaxRegs];  // LDS byte offsets of the incoming transpose's writes
    uint32_t dst_lswz[kMaxRegs];  // ... and of its reads (lay.lswz unless split)
    uint32_t op_begin, op_end;
    uint32_t transpose_in;  // 1: go through LDS from the previous layout first
    uint32_t pad;
};

// One HBM pass: a tile is the 2^k amplitudes whose index bits outside
// qpos[0..k) are fixed by the workgroup id.
struct alignas(16) DevPass {
    uint64_t greg_ld[kMaxRegs];  // global offset of register r when loading
    uint64_t greg_st[kMaxRegs];  // ... when storing
    uint32_t gthr_ld[kMaxThreadBits];  // global bit of thread bit j when loading
    uint32_t gthr_st[kMaxThreadBits];
    uint32_t qpos[16];           // global qubit of local bit j, ascending
    DevLayout lay_ld, lay_st;
    uint32_t st_src_lswz[kMaxRegs];  // LDS byte offsets of the store transpose's writes
    uint32_t st_dst_lswz[kMaxRegs];  // ... and of its reads
    uint32_t phase_begin, phase_end;
    uint32_t transpose_st;       // 1: go through LDS to lay_st before storing
    uint32_t k;
    // Split transposes: one local bit w sits at the top thread bit of every
    // layout of the pass, so the threads holding w = 0 and those holding
    // w = 1 exchange amplitudes only among themselves.  Every transpose of
    // the pass then runs as two rounds through half the tile's LDS, and its
    // LDS byte offsets drop bit `split` (= w + log2 of the amplitude size,
    // kept by the swizzle since w >= its group width).  0: no split.
    uint32_t split;
    uint32_t lds_half;  // 1: the pass fits kernels with half a tile of LDS
    uint32_t pad[2];
};

// A diagonal the kernels apply as sign flips: exactly diag(1, -1).
inline bool diag_is_sign_flip(const double *m)
{
    return m[0] == 1.0 && m[1] == 0.0 && m[2] == 0.0 && m[3] == 0.0 && m[4] == 0.0 && m[5] == 0.0 &&
           m[6] == -1.0 && m[7] == 0.0;
}

// An LDS byte offset with bit b removed (the bits above it shift down).
inline DMV_HD uint32_t drop_bit(uint32_t y, uint32_t b)
{
    const uint32_t lo = (1u << b) - 1;
    return (y & lo) | ((y >> 1) & ~lo);
}

// ---------------------------------------------------------- bit helpers --
inline DMV_HD uint64_t insert_zero_bit(uint64_t x, uint32_t pos)
{
    const uint64_t low = x & ((uint64_t(1) << pos) - 1);
    return ((x >> pos) << (pos + 1)) | low;
}

inline int popcount64(uint64_t x) { return __builtin_popcountll(x); }

// ------------------------------------------------------- state layout --
// Amplitude i of a state lives at element
//     i ^ (P(s) << blk) ^ (F(s >> 6) << shift),    s = i >> shift.
// The state is cut into 2 MiB spans of 1 KiB blocks.
// - Within span s the block index is XORed with P(s), P the GF(2)-linear map
//   sending bit j of the span index to bit (7 j mod 11) of the block index.
//   In the plain order, the amplitudes a fused pass gathers for a tile over
//   high qubits sit at large power-of-two distances and land on the same HBM
//   channels; the swizzle spreads them (measured on tile-shaped copies at 30
//   qubits: 3.9-4.1 TB/s plain against 5.1-5.4 TB/s swizzled, with contiguous
//   tiles unchanged).
// - The span index itself has its bits 6 and up (spans 128 MiB apart) folded
//   onto its top `fold` low bits: F sends bit j to bit (6 - fold + j mod fold).
//   A tile whose high qubits all lie 128 MiB or more apart otherwise touches
//   only spans equal in their six low bits; on tile-shaped copies at 33
//   qubits, qubits 23-31 ran at 2.4 TB/s plain and 5.9 TB/s folded.
// The map is a bijection (no extra memory), keeps 1 KiB blocks contiguous and
// is linear: elem(a ^ b) = elem(a) ^ elem(b), so kernels combine a tile base,
// a thread offset and register offsets with XOR.  States of at most one span
// have the plain layout.
struct StateLayout {
    uint32_t shift;  // log2 amplitudes per span (0: plain layout)
    uint32_t blk;    // log2 amplitudes per block
    uint32_t fold;   // low span-index bits that the high ones fold onto (0: none)
    uint32_t pad;
};

constexpr uint32_t kFoldBase = 6;  // span-index bits at and above this fold down

// The layout of an n-qubit state with the given fold width (a state records
// its own layout when it is created).
inline StateLayout state_layout(int n, bool c64, uint32_t fold)
{
    const uint32_t shift = c64 ? 18u : 17u;  // 2 MiB of amplitudes
    const uint32_t blk = c64 ? 7u : 6u;      // 1 KiB
    if (n <= (int)shift)
        return StateLayout{0u, 0u, 0u, 0u};
    return StateLayout{shift, blk, fold > kFoldBase ? kFoldBase : fold, 0u};
}

inline bool same_layout(const StateLayout &a, const StateLayout &b)
{
    return a.shift == b.shift && a.blk == b.blk && a.fold == b.fold;
}

// P: bit j of the span index -> bit (7 j mod 11) of the block index.
inline DMV_HD uint64_t span_swizzle(uint64_t s)
{
    uint64_t o = 0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int j = 0; j < 16; j++)
        o ^= ((s >> j) & 1u) << ((7 * j) % 11);
    return o;
}

// F: bit j of h -> bit (kFoldBase - f + j mod f), f in [1, kFoldBase].
inline DMV_HD uint64_t span_fold(uint64_t h, uint32_t f)
{
    uint64_t o = 0;
    for (int k = 0; k < 16 && h; k++) {
        o ^= h;
        h >>= f;
    }
    return (o & ((uint64_t(1) << f) - 1)) << (kFoldBase - f);
}

inline DMV_HD uint64_t elem(uint64_t i, StateLayout L)
{
    if (!L.shift)
        return i;
    const uint64_t s = i >> L.shift;
    uint64_t e = i ^ (span_swizzle(s) << L.blk);
    if (L.fold)
        e ^= span_fold(s >> kFoldBase, L.fold) << L.shift;
    return e;
}

static_assert(kSmallMaxQubitsC128 <= 17 && kSmallMaxQubitsC64 <= 18,
              "whole-state kernels index states of the plain layout");
static_assert(kMaxQubits - 17 <= 16, "span index wider than the swizzle");

// Layout of private scratch states (batched circuits).
constexpr StateLayout kPlainLayout{0u, 0u, 0u, 0u};

}  // namespace dmv
