// Device building blocks of the fused tile executor, shared by the big-state
// pass kernels, the whole-state kernels and the batched persistent kernel.
//
// The host resolves every op against the phase's register/thread layout and
// against the register renaming in force (see make_pass in plan.cpp):
//  * pair ops touch physical registers p and p ^ rmask; the pair loop is
//    generated per mask value, so register indices stay compile-time;
//  * which member of a pair is the target-bit-0 one (fl), which registers
//    pass their register-bit controls (en) and which registers of a
//    register-target diagonal have target bit 1 (rmask) are uniform masks,
//    so choosing coefficients costs scalar selects only;
//  * thread-bit controls are one predicate per op.
// The vector ALU therefore does only the complex arithmetic.
#pragma once

#include "device.hpp"

namespace dmv {

template <int TLOG> __device__ __forceinline__ uint32_t thread_local_offset(const DevLayout *L, uint32_t tid)
{
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < TLOG; j++)
        o |= ((tid >> j) & 1u) << L->thr[j];
    return o;
}

template <int TLOG> __device__ __forceinline__ uint64_t thread_global_offset(const uint32_t *gthr, uint32_t tid)
{
    uint64_t o = 0;
#pragma unroll
    for (int j = 0; j < TLOG; j++)
        o |= (uint64_t)((tid >> j) & 1u) << gthr[j];
    return o;
}

// Moves the tile through LDS: register p of thread t is written at
// swz(thread part in A) ^ src[p] and register r is read back from
// swz(thread part in B) ^ dst[r].  The swizzle is linear, so these are
// the swizzled local indices of (t, p) in A and (t, r) in B.
template <class V, int NR, int TLOG>
__device__ __forceinline__ void transpose(V (&v)[NR], V *lds, const DevLayout *A, const uint32_t *src,
                                          const DevLayout *B, const uint32_t *dst, uint32_t tid)
{
    constexpr int G = sizeof(V) == 16 ? 3 : 4;
    char *base = reinterpret_cast<char *>(lds);
    const uint32_t la = swz_index<G>(thread_local_offset<TLOG>(A, tid)) * (uint32_t)sizeof(V);
#pragma unroll
    for (int r = 0; r < NR; r++)
        *reinterpret_cast<V *>(base + (la ^ src[r])) = v[r];
    __syncthreads();
    const uint32_t lb = swz_index<G>(thread_local_offset<TLOG>(B, tid)) * (uint32_t)sizeof(V);
#pragma unroll
    for (int r = 0; r < NR; r++)
        v[r] = *reinterpret_cast<const V *>(base + (lb ^ dst[r]));
    __syncthreads();
}

// The same through half a tile of LDS, in a pass whose layouts all hold one
// local bit w at the top thread bit (a wave bit): the threads of w = 0
// exchange their amplitudes first, then those of w = 1.  Offsets drop byte
// bit `split` (w's bit, which the swizzle leaves in place); src and dst come
// from the host with that bit already removed, and removing a bit is linear,
// so drop(la ^ src[p]) = drop(la) ^ src[p].  Every wave passes the same four
// barriers; the rounds are uniform branches on the wave's half.
template <class V, int NR, int TLOG>
__device__ __forceinline__ void transpose_half(V (&v)[NR], V *lds, const DevLayout *A, const uint32_t *src,
                                               const DevLayout *B, const uint32_t *dst, uint32_t split,
                                               uint32_t tid)
{
    static_assert(TLOG >= 7, "the split bit must be a wave bit");
    constexpr int G = sizeof(V) == 16 ? 3 : 4;
    char *base = reinterpret_cast<char *>(lds);
    const uint32_t la =
        drop_bit(swz_index<G>(thread_local_offset<TLOG>(A, tid)) * (uint32_t)sizeof(V), split);
    const uint32_t lb =
        drop_bit(swz_index<G>(thread_local_offset<TLOG>(B, tid)) * (uint32_t)sizeof(V), split);
    const uint32_t half = __builtin_amdgcn_readfirstlane(tid >> (TLOG - 1));
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        if (half == h) {
#pragma unroll
            for (int r = 0; r < NR; r++)
                *reinterpret_cast<V *>(base + (la ^ src[r])) = v[r];
        }
        __syncthreads();
        if (half == h) {
#pragma unroll
            for (int r = 0; r < NR; r++)
                v[r] = *reinterpret_cast<const V *>(base + (lb ^ dst[r]));
        }
        __syncthreads();
    }
}

constexpr int highest_bit(int m) { return m >= 16 ? 16 : m >= 8 ? 8 : m >= 4 ? 4 : m >= 2 ? 2 : 1; }

// In-place pair ops of the reverse sweep.  A GEN (plain or not: one code
// path keeps the kernel small, and a plain GEN's selects all pick A) on
// registers (p, p ^ M) of psi and lambda at once, so the per-pair
// coefficients (scalar selects) are chosen once for both states.  f: p
// holds the target-bit-1 member, so the matrix is applied with rows and
// columns exchanged.  GEN ops carry no controls (plans only make them from
// uncontrolled gates, and plan validation checks it), so no pair or lane is
// left out.
template <int M, class V, class R, int NR>
__device__ __forceinline__ void op_gen2_m(V (&a)[NR], V (&b)[NR], const Mat<R> &A, uint32_t fl)
{
    constexpr int HB = highest_bit(M);
#pragma unroll
    for (int p = 0; p < NR; p++) {
        if (p & HB)
            continue;
        const bool f = (fl >> p) & 1u;
        Mat<R> B;
        B.m00r = f ? A.m11r : A.m00r; B.m00i = f ? A.m11i : A.m00i;
        B.m01r = f ? A.m10r : A.m01r; B.m01i = f ? A.m10i : A.m01i;
        B.m10r = f ? A.m01r : A.m10r; B.m10i = f ? A.m01i : A.m10i;
        B.m11r = f ? A.m00r : A.m11r; B.m11i = f ? A.m00i : A.m11i;
        apply2(a[p], a[p ^ M], B);
        __builtin_amdgcn_sched_barrier(0);
        apply2(b[p], b[p ^ M], B);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Swaps and diagonals act where the pair's register controls (en) and the
// lane's thread controls (hold) pass; both enter the selects they make
// anyway, so that no lane branches around an op (a divergent branch around
// in-place updates makes the compiler copy every register of both states
// before each op).
template <int M, class V, int NR> __device__ __forceinline__ void op_swap_m(V (&v)[NR], uint32_t en, bool hold)
{
    constexpr int HB = highest_bit(M);
#pragma unroll
    for (int p = 0; p < NR; p++) {
        if (p & HB)
            continue;
        const bool ok = ((en >> p) & 1u) && hold;
        const V x = v[p], y = v[p ^ M];
        v[p] = ok ? y : x;
        v[p ^ M] = ok ? x : y;
    }
}

#define DMV_MASK_CASES(X)                                                                          \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17)   \
    X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)
// The same without 1: the out-of-place switches send mask 1 and anything out
// of range (which the host's plan validation excludes) to one case body, so
// that every path writes its outputs in a case.
#define DMV_MASK_CASES_FROM2(X)                                                                    \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17)        \
    X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)

// A GEN on psi and lambda (one dispatch, shared coefficient selects).
template <class V, class R, int NR>
__device__ __forceinline__ void op_gen2(V (&a)[NR], V (&b)[NR], const DevOp *op)
{
    const uint32_t fl = op->fl;
    const Mat<R> A = load_mat<R>(op->m);
    switch (op->rmask) {
#define DMV_GEN2_CASE(M)                                                                           \
    case M:                                                                                        \
        if constexpr (M < NR)                                                                      \
            op_gen2_m<M>(a, b, A, fl);                                                             \
        break;
        DMV_MASK_CASES(DMV_GEN2_CASE)
#undef DMV_GEN2_CASE
    default: break;
    }
}

// SWAP (X / CNOT that renaming could not absorb) on registers (p, p ^ rmask).
template <class V, int NR>
__device__ __forceinline__ void op_swap(V (&v)[NR], const DevOp *op, bool hold)
{
    const uint32_t en = op->en;
    switch (op->rmask) {
#define DMV_SWAP_CASE(M)                                                                           \
    case M:                                                                                        \
        if constexpr (M < NR)                                                                      \
            op_swap_m<M>(v, en, hold);                                                             \
        break;
        DMV_MASK_CASES(DMV_SWAP_CASE)
#undef DMV_SWAP_CASE
    default: break;
    }
}

// v with the signs of both parts flipped where m is the sign bit (m = 0
// leaves it as it is): -v exactly, by integer XOR.
template <class V> __device__ __forceinline__ V flip_sign(V v, uint32_t m)
{
    if constexpr (sizeof(V) == 16) {
        using U = uint32_t __attribute__((ext_vector_type(4)));
        U u = __builtin_bit_cast(U, v);
        u.y ^= m;
        u.w ^= m;
        return __builtin_bit_cast(V, u);
    } else {
        using U = uint32_t __attribute__((ext_vector_type(2)));
        U u = __builtin_bit_cast(U, v);
        u.x ^= m;
        u.y ^= m;
        return __builtin_bit_cast(V, u);
    }
}

// The sign mask of register r for a diag(1, -1) op (plain DIAG): the sign bit
// where the register holds target bit 1 and the register and thread
// controls hold.  lane: the sign bit in lanes whose thread controls hold
// (and, for a thread-bit target, whose target bit is 1).
template <int NR>
__device__ __forceinline__ uint32_t flip_mask(const DevOp *op, int r, uint32_t lane)
{
    const uint32_t all = NR >= 32 ? 0xFFFFFFFFu : ((1u << NR) - 1);
    const bool en_r = op->en == all || ((op->en >> r) & 1u);
    const bool one_r = op->tbit != kNoReg || ((op->rmask >> r) & 1u);
    return en_r && one_r ? lane : 0u;
}

__device__ __forceinline__ uint32_t flip_lane(const DevOp *op, uint32_t tid, bool hold)
{
    const bool one = op->tbit == kNoReg || ((tid >> op->tbit) & 1u);
    return hold && one ? 0x80000000u : 0u;
}

// diag(m00, m11) on a register-bit target (rmask: registers with target bit
// 1) or on a thread-bit target (tbit), on the registers in en, in lanes
// whose thread controls hold.  diag(1, -1) (plain) flips signs instead of
// multiplying.
template <class V, class R, int NR>
__device__ __forceinline__ void op_diag(V (&v)[NR], const DevOp *op, uint32_t tid, bool hold)
{
    if (op->plain) {
        const uint32_t lane = flip_lane(op, tid, hold);
#pragma unroll
        for (int r = 0; r < NR; r++)
            v[r] = flip_sign(v[r], flip_mask<NR>(op, r, lane));
        return;
    }
    const R d0r = (R)op->m[0], d0i = (R)op->m[1], d1r = (R)op->m[6], d1i = (R)op->m[7];
    const uint32_t en = op->en;
    const uint32_t all = NR >= 32 ? 0xFFFFFFFFu : ((1u << NR) - 1);
    if (op->tbit == kNoReg) {
        const uint32_t ones = op->rmask;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const bool one = (ones >> r) & 1u;
            const V x = cmul(v[r], one ? d1r : d0r, one ? d1i : d0i);
            v[r] = ((en == all || ((en >> r) & 1u)) && hold) ? x : v[r];
        }
    } else {
        const bool one = (tid >> op->tbit) & 1u;
        const R dr = one ? d1r : d0r, di = one ? d1i : d0i;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const V x = cmul(v[r], dr, di);
            v[r] = ((en == all || ((en >> r) & 1u)) && hold) ? x : v[r];
        }
    }
}

__device__ __forceinline__ bool thread_controls_hold(const DevOp *op, uint32_t tid)
{
    const uint32_t ct = op->ctrl_thr;
    return (tid & ct) == ct;
}

// ------------------------------------------------ out-of-place op bodies --
// Each writes every element of `out` from `in`.  An op loop that alternates
// two register arrays then merges the cases of an op's switch without
// copies: with an in-place loop the compiler keeps the loop-carried array in
// one register set and the cases' results in another, and moves all 2^RLOG
// amplitudes between them twice per op (measured: 64 v_mov_b64 per op of the
// complex128 k = 12 kernel, a third of its VALU instructions).

template <class V, class R>
__device__ __forceinline__ void apply2_to(V &oa, V &ob, const V &a, const V &b, const Mat<R> &M)
{
    oa.x = M.m00r * a.x - M.m00i * a.y + M.m01r * b.x - M.m01i * b.y;
    oa.y = M.m00r * a.y + M.m00i * a.x + M.m01r * b.y + M.m01i * b.x;
    ob.x = M.m10r * a.x - M.m10i * a.y + M.m11r * b.x - M.m11i * b.y;
    ob.y = M.m10r * a.y + M.m10i * a.x + M.m11r * b.y + M.m11i * b.x;
}

template <class V, int NR> __device__ __forceinline__ void copy_regs(V (&out)[NR], const V (&in)[NR])
{
#pragma unroll
    for (int r = 0; r < NR; r++)
        out[r] = in[r];
}

template <int M, class V, class R, int NR>
__device__ __forceinline__ void oop_gen_plain_m(V (&out)[NR], const V (&in)[NR], const Mat<R> &A)
{
    constexpr int HB = highest_bit(M);
#pragma unroll
    for (int p = 0; p < NR; p++) {
        if (p & HB)
            continue;
        apply2_to(out[p], out[p ^ M], in[p], in[p ^ M], A);
    }
}

// A GEN whose pair members are swapped where fl says so (GENs carry no
// controls, so every pair of every lane takes it).
template <int M, class V, class R, int NR>
__device__ __forceinline__ void oop_gen_m(V (&out)[NR], const V (&in)[NR], const Mat<R> &A, uint32_t fl)
{
    constexpr int HB = highest_bit(M);
#pragma unroll
    for (int p = 0; p < NR; p++) {
        if (p & HB)
            continue;
        const bool f = (fl >> p) & 1u;
        Mat<R> B;
        B.m00r = f ? A.m11r : A.m00r; B.m00i = f ? A.m11i : A.m00i;
        B.m01r = f ? A.m10r : A.m01r; B.m01i = f ? A.m10i : A.m01i;
        B.m10r = f ? A.m01r : A.m10r; B.m10i = f ? A.m01i : A.m10i;
        B.m11r = f ? A.m00r : A.m11r; B.m11i = f ? A.m00i : A.m11i;
        apply2_to(out[p], out[p ^ M], in[p], in[p ^ M], B);
    }
}

template <int M, class V, int NR>
__device__ __forceinline__ void oop_swap_m(V (&out)[NR], const V (&in)[NR], uint32_t en, bool hold)
{
    constexpr int HB = highest_bit(M);
#pragma unroll
    for (int p = 0; p < NR; p++) {
        if (p & HB)
            continue;
        const bool ok = ((en >> p) & 1u) && hold;
        out[p] = ok ? in[p ^ M] : in[p];
        out[p ^ M] = ok ? in[p] : in[p ^ M];
    }
}

template <class V, class R, int NR>
__device__ __forceinline__ void oop_diag(V (&out)[NR], const V (&in)[NR], const DevOp *op, uint32_t tid, bool hold)
{
    if (op->plain) {  // diag(1, -1): sign flips
        const uint32_t lane = flip_lane(op, tid, hold);
#pragma unroll
        for (int r = 0; r < NR; r++)
            out[r] = flip_sign(in[r], flip_mask<NR>(op, r, lane));
        return;
    }
    const R d0r = (R)op->m[0], d0i = (R)op->m[1], d1r = (R)op->m[6], d1i = (R)op->m[7];
    const uint32_t en = op->en;
    const uint32_t all = NR >= 32 ? 0xFFFFFFFFu : ((1u << NR) - 1);
    if (op->tbit == kNoReg) {
        const uint32_t ones = op->rmask;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const bool one = (ones >> r) & 1u;
            const V x = cmul(in[r], one ? d1r : d0r, one ? d1i : d0i);
            out[r] = ((en == all || ((en >> r) & 1u)) && hold) ? x : in[r];
        }
    } else {
        const bool one = (tid >> op->tbit) & 1u;
        const R dr = one ? d1r : d0r, di = one ? d1i : d0i;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const V x = cmul(in[r], dr, di);
            out[r] = ((en == all || ((en >> r) & 1u)) && hold) ? x : in[r];
        }
    }
}

// One op from `in` into `out`; lanes whose thread-bit controls do not hold
// keep their amplitudes through the selects that swaps and diagonals make
// anyway (GENs carry no controls).
template <class V, class R, int NR>
__device__ __forceinline__ void oop_op(V (&out)[NR], const V (&in)[NR], const DevOp *op, uint32_t tid)
{
    const uint32_t en = op->en, fl = op->fl;
    const bool hold = thread_controls_hold(op, tid);
    const Mat<R> A = load_mat<R>(op->m);
    // masks at or above NR have empty cases; the clamp lets the compiler drop them
    const uint32_t m = op->rmask < (uint32_t)NR ? op->rmask : 1u;
    if (op->kind == OP_DIAG) {
        oop_diag<V, R>(out, in, op, tid, hold);
    } else if (op->plain) {
        switch (m) {
#define DMV_OOP_PLAIN_CASE(M)                                                                      \
    case M:                                                                                        \
        if constexpr (M < NR)                                                                      \
            oop_gen_plain_m<M>(out, in, A);                                                        \
        break;
            DMV_MASK_CASES_FROM2(DMV_OOP_PLAIN_CASE)
#undef DMV_OOP_PLAIN_CASE
        default: oop_gen_plain_m<1>(out, in, A); break;
        }
    } else if (op->kind == OP_GEN) {
        switch (m) {
#define DMV_OOP_GEN_CASE(M)                                                                        \
    case M:                                                                                        \
        if constexpr (M < NR)                                                                      \
            oop_gen_m<M>(out, in, A, fl);                                                          \
        break;
            DMV_MASK_CASES_FROM2(DMV_OOP_GEN_CASE)
#undef DMV_OOP_GEN_CASE
        default: oop_gen_m<1>(out, in, A, fl); break;
        }
    } else {
        switch (m) {
#define DMV_OOP_SWAP_CASE(M)                                                                       \
    case M:                                                                                        \
        if constexpr (M < NR)                                                                      \
            oop_swap_m<M>(out, in, en, hold);                                                      \
        break;
            DMV_MASK_CASES_FROM2(DMV_OOP_SWAP_CASE)
#undef DMV_OOP_SWAP_CASE
        default: oop_swap_m<1>(out, in, en, hold); break;
        }
    }
}

// The ops [ob, oe) of a phase, alternating between v and a second array.
// Each op is copied whole into scalar registers first: a few wide scalar
// loads and one wait, instead of a load and a wait per field as the code
// reaches it.
template <class V, class R, int RLOG>
__device__ __forceinline__ void run_ops(V (&v)[1 << RLOG], const DevOp *__restrict__ ops, uint32_t ob, uint32_t oe,
                                        uint32_t tid)
{
    V w[1 << RLOG];
    uint32_t o = ob;
    for (; o + 1 < oe; o += 2) {
        const DevOp a = ops[o];
        oop_op<V, R>(w, v, &a, tid);
        const DevOp b = ops[o + 1];
        oop_op<V, R>(v, w, &b, tid);
    }
    if (o < oe) {
        const DevOp a = ops[o];
        oop_op<V, R>(w, v, &a, tid);
        copy_regs(v, w);
    }
}

// Global index of a tile's first amplitude: the tile number with a zero bit
// inserted at every local qubit position.
template <int K> __device__ __forceinline__ uint64_t tile_base(const DevPass *P, uint64_t tile)
{
    uint64_t g = tile;
#pragma unroll
    for (int j = 0; j < K; j++)
        g = insert_zero_bit(g, P->qpos[j]);
    return g;
}

// g (tile base and thread offset) and the register offsets have disjoint
// bits, so the element of amplitude g | reg is elem(g) ^ elem(reg); greg holds
// the register offsets already mapped to elements by the planner.
template <bool NT, class V, int NR>
__device__ __forceinline__ void load_tile(V (&v)[NR], const V *__restrict__ state, uint64_t g, const uint64_t *greg,
                                          StateLayout L)
{
    const uint64_t eg = elem(g, L);
#pragma unroll
    for (int r = 0; r < NR; r++)
        v[r] = ld<NT>(state + (eg ^ greg[r]));
}

template <bool NT, class V, int NR>
__device__ __forceinline__ void store_tile(const V (&v)[NR], V *__restrict__ state, uint64_t g, const uint64_t *greg,
                                           StateLayout L)
{
    const uint64_t eg = elem(g, L);
#pragma unroll
    for (int r = 0; r < NR; r++)
        st<NT>(state + (eg ^ greg[r]), v[r]);
}

// Runs the phases of a pass on a loaded tile and leaves it in the store
// layout.  HALF: lds holds half a tile and every transpose of the pass is
// split (the host runs only such passes, or passes without transposes, on
// these kernels).
template <class R, int RLOG, int TLOG, bool HALF>
__device__ __forceinline__ void process_tile(typename Cplx<R>::T (&v)[1 << RLOG], const DevPass *P,
                                             const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops,
                                             typename Cplx<R>::T *lds)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG;
    const uint32_t tid = threadIdx.x;
    const DevLayout *cur = &P->lay_ld;
    for (uint32_t ph = P->phase_begin; ph < P->phase_end; ph++) {
        const DevPhase *D = phases + ph;
        if (D->transpose_in) {
            if constexpr (HALF)
                transpose_half<V, NR, TLOG>(v, lds, cur, D->src_lswz, &D->lay, D->dst_lswz, P->split, tid);
            else
                transpose<V, NR, TLOG>(v, lds, cur, D->src_lswz, &D->lay, D->dst_lswz, tid);
        }
        cur = &D->lay;
        run_ops<V, R, RLOG>(v, ops, D->op_begin, D->op_end, tid);
    }
    if (P->transpose_st) {
        if constexpr (HALF)
            transpose_half<V, NR, TLOG>(v, lds, cur, P->st_src_lswz, &P->lay_st, P->st_dst_lswz, P->split, tid);
        else
            transpose<V, NR, TLOG>(v, lds, cur, P->st_src_lswz, &P->lay_st, P->st_dst_lswz, tid);
    }
}

// One tile of a forward pass: load, run the phases, store.
template <class R, int RLOG, int TLOG, bool NT, bool HALF = false>
__device__ __forceinline__ void tile_body(typename Cplx<R>::T *__restrict__ state, const DevPass *P,
                                          const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops,
                                          uint64_t tile, typename Cplx<R>::T *lds, StateLayout L)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG, K = RLOG + TLOG;
    const uint32_t tid = threadIdx.x;
    const uint64_t g = tile_base<K>(P, tile);
    V v[NR];
    load_tile<NT>(v, state, g | thread_global_offset<TLOG>(P->gthr_ld, tid), P->greg_ld, L);
    process_tile<R, RLOG, TLOG, HALF>(v, P, phases, ops, lds);
    store_tile<NT>(v, state, g | thread_global_offset<TLOG>(P->gthr_st, tid), P->greg_st, L);
}

// --------------------------------------------------------------- adjoint --
// Per-thread sums for an IP op on a register-bit target: over register pairs
// (a, a ^ M) the parts of Im<lambda|X|psi>, Im<lambda|Y|psi> and
// Im<lambda|Z|psi>.  With u0, u1 (v0, v1) the psi (lambda) amplitudes of
// target bit 0 and 1: X gives Im(v0* u1) + Im(v1* u0), Y gives
// -Re(v0* u1) + Re(v1* u0) and Z gives Im(v0* u0) - Im(v1* u1); the first
// member a holds bit 1 where fl says so, which flips the signs of Y and Z.
template <int M, class V, int NR>
__device__ __forceinline__ void ip_pair_m(const V (&p)[NR], const V (&l)[NR], uint32_t fl, double &sx, double &sy,
                                          double &sz)
{
    constexpr int HB = highest_bit(M);
#pragma unroll
    for (int a = 0; a < NR; a++) {
        if (a & HB)
            continue;
        const double pax = p[a].x, pay = p[a].y, pbx = p[a ^ M].x, pby = p[a ^ M].y;
        const double lax = l[a].x, lay = l[a].y, lbx = l[a ^ M].x, lby = l[a ^ M].y;
        sx = __builtin_fma(lax, pby, sx);
        sx = __builtin_fma(-lay, pbx, sx);
        sx = __builtin_fma(lbx, pay, sx);
        sx = __builtin_fma(-lby, pax, sx);
        double y = lbx * pax;
        y = __builtin_fma(lby, pay, y);
        y = __builtin_fma(-lax, pbx, y);
        y = __builtin_fma(-lay, pby, y);
        double z = lax * pay;
        z = __builtin_fma(-lay, pax, z);
        z = __builtin_fma(-lbx, pby, z);
        z = __builtin_fma(lby, pbx, z);
        const double sg = ((fl >> a) & 1u) ? -1.0 : 1.0;  // a scalar: fl is uniform
        sy = __builtin_fma(sg, y, sy);
        sz = __builtin_fma(sg, z, sz);
    }
}

// This thread's sums for an IP op: (sx, sy, sz) for IPG; sz alone for IPZ,
// whose target may be a thread bit.
template <class V, int RLOG>
__device__ __forceinline__ void ip_sums(const V (&p)[1 << RLOG], const V (&l)[1 << RLOG], const DevOp *op,
                                        uint32_t tid, double &sx, double &sy, double &sz)
{
    constexpr int NR = 1 << RLOG;
    sx = sy = sz = 0.0;
    if (op->kind == OP_IPZ) {
        const bool reg = op->tbit == kNoReg;
        const bool thread_one = !reg && ((tid >> op->tbit) & 1u);
        const uint32_t ones = op->rmask;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const bool one = reg ? ((ones >> r) & 1u) : thread_one;
            double w = (double)l[r].x * p[r].y;
            w = __builtin_fma(-(double)l[r].y, p[r].x, w);
            sz = __builtin_fma(one ? -1.0 : 1.0, w, sz);
        }
        return;
    }
    switch (op->rmask) {
#define DMV_IP_CASE(M)                                                                             \
    case M:                                                                                        \
        if constexpr (M < NR)                                                                      \
            ip_pair_m<M>(p, l, op->fl, sx, sy, sz);                                                \
        break;
        DMV_MASK_CASES(DMV_IP_CASE)
#undef DMV_IP_CASE
    default: break;
    }
}

// One tile of a reverse-sweep pass over psi and lambda.  For every IP op the
// sum over each wave (Sink::kPerWave) or over the workgroup is handed to
// sink(op, value, wave) by the first lane of the wave / thread 0.  The ops
// are applied in place, to psi and then to lambda: an out-of-place loop over
// two states (four register arrays, or three with one run of ops per state
// between IP ops) needed more registers or twice the code and ran slower.
// STREAM: streaming (nontemporal) loads and stores of both states.
template <class R, int RLOG, int TLOG, bool STREAM = false, class Sink>
__device__ __forceinline__ void tile_body_adjoint(typename Cplx<R>::T *__restrict__ psi,
                                                  typename Cplx<R>::T *__restrict__ lam, const DevPass *P,
                                                  const DevPhase *__restrict__ phases,
                                                  const DevOp *__restrict__ ops, uint64_t tile,
                                                  typename Cplx<R>::T *lds, double *red, Sink sink, StateLayout L)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG, K = RLOG + TLOG, NT = 1 << TLOG;
    const uint32_t tid = threadIdx.x;
    const uint64_t g = tile_base<K>(P, tile);
    const uint64_t gl = g | thread_global_offset<TLOG>(P->gthr_ld, tid);
    V p[NR], l[NR];
    load_tile<STREAM>(p, psi, gl, P->greg_ld, L);
    load_tile<STREAM>(l, lam, gl, P->greg_ld, L);
    const DevLayout *cur = &P->lay_ld;
    double sx = 0.0, sy = 0.0, sz = 0.0;  // sums of the current IP group
    for (uint32_t ph = P->phase_begin; ph < P->phase_end; ph++) {
        const DevPhase *D = phases + ph;
        if (D->transpose_in) {
            transpose<V, NR, TLOG>(p, lds, cur, D->src_lswz, &D->lay, D->dst_lswz, tid);
            transpose<V, NR, TLOG>(l, lds, cur, D->src_lswz, &D->lay, D->dst_lswz, tid);
        }
        cur = &D->lay;
        for (uint32_t o = D->op_begin; o < D->op_end; o++) {
            // the whole op in scalar registers at once (a few wide scalar
            // loads and one wait instead of one load and wait per field)
            const DevOp opv = ops[o];
            const DevOp *op = &opv;
            if (op_is_ip(op->kind)) {
                if (!op->plain)
                    ip_sums<V, RLOG>(p, l, op, tid, sx, sy, sz);
                const double v = op->m[0] * sx + op->m[1] * sy + op->m[2] * sz;
                if constexpr (Sink::kPerWave) {
                    // one partial per wave: lane moves only, no workgroup barrier
                    const double w = wave_sum_rows(v);
                    if ((tid & 63) == 0)
                        sink(op, w, tid >> 6);
                } else {
                    const double s = block_sum<NT>(v, red);
                    if (tid == 0)
                        sink(op, s, 0);
                }
                continue;
            }
            const bool hold = thread_controls_hold(op, tid);
            if (op->kind == OP_DIAG) {
                op_diag<V, R>(p, op, tid, hold);
                op_diag<V, R>(l, op, tid, hold);
            } else if (op->kind == OP_GEN) {
                op_gen2<V, R>(p, l, op);
            } else {
                op_swap<V>(p, op, hold);
                op_swap<V>(l, op, hold);
            }
        }
    }
    if (P->transpose_st) {
        transpose<V, NR, TLOG>(p, lds, cur, P->st_src_lswz, &P->lay_st, P->st_dst_lswz, tid);
        transpose<V, NR, TLOG>(l, lds, cur, P->st_src_lswz, &P->lay_st, P->st_dst_lswz, tid);
    }
    const uint64_t gs = g | thread_global_offset<TLOG>(P->gthr_st, tid);
    store_tile<STREAM>(p, psi, gs, P->greg_st, L);
    store_tile<STREAM>(l, lam, gs, P->greg_st, L);
}

// ------------------------------------------------------ whole-state in LDS --
// One op over a state of 2^n amplitudes held in LDS by the whole workgroup.
template <class V, class R>
__device__ __forceinline__ void small_apply(V *s, uint32_t n, const DevOp *op)
{
    const uint32_t n_amp = 1u << n;
    const uint32_t t = op->target;
    const uint32_t ctrl = (uint32_t)op->ctrl;
    if (op->kind == OP_DIAG) {
        const R d0r = (R)op->m[0], d0i = (R)op->m[1], d1r = (R)op->m[6], d1i = (R)op->m[7];
        for (uint32_t i = threadIdx.x; i < n_amp; i += blockDim.x)
            if ((i & ctrl) == ctrl) {
                const bool one = (i >> t) & 1u;
                s[i] = cmul(s[i], one ? d1r : d0r, one ? d1i : d0i);
            }
    } else if (op->kind == OP_SWAP) {
        for (uint32_t p = threadIdx.x; p < n_amp / 2; p += blockDim.x) {
            const uint32_t i0 = (uint32_t)insert_zero_bit(p, t), i1 = i0 | (1u << t);
            if ((i0 & ctrl) == ctrl) {
                const V a = s[i0];
                s[i0] = s[i1];
                s[i1] = a;
            }
        }
    } else if (op->kind == OP_GEN) {
        const Mat<R> M = load_mat<R>(op->m);
        for (uint32_t p = threadIdx.x; p < n_amp / 2; p += blockDim.x) {
            const uint32_t i0 = (uint32_t)insert_zero_bit(p, t), i1 = i0 | (1u << t);
            if ((i0 & ctrl) == ctrl) {
                V a = s[i0], b = s[i1];
                apply2(a, b, M);
                s[i0] = a;
                s[i1] = b;
            }
        }
    }
}

// This thread's part of Im<lambda|A|psi>, A = ax X + ay Y + az Z on the
// target, for an IP op on a whole LDS state.
template <class V>
__device__ __forceinline__ double small_ip(const V *sp, const V *sl, uint32_t n, const DevOp *op)
{
    const uint32_t n_amp = 1u << n, t = op->target;
    const double ax = op->m[0], ay = op->m[1], az = op->m[2];
    double acc = 0.0;
    for (uint32_t q = threadIdx.x; q < n_amp / 2; q += blockDim.x) {
        const uint32_t i0 = (uint32_t)insert_zero_bit(q, t), i1 = i0 | (1u << t);
        const V u0 = sp[i0], u1 = sp[i1], v0 = sl[i0], v1 = sl[i1];
        const double x = (double)v0.x * u1.y - (double)v0.y * u1.x + (double)v1.x * u0.y - (double)v1.y * u0.x;
        const double y = -((double)v0.x * u1.x + (double)v0.y * u1.y) + ((double)v1.x * u0.x + (double)v1.y * u0.y);
        const double z = ((double)v0.x * u0.y - (double)v0.y * u0.x) - ((double)v1.x * u1.y - (double)v1.y * u1.x);
        acc += ax * x + ay * y + az * z;
    }
    return acc;
}

}  // namespace dmv
