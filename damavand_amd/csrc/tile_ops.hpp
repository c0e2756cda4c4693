// Device building blocks of the fused tile executor, shared by the big-state
// pass kernels, the whole-state kernels and the batched persistent kernel.
//
// Every per-op condition is resolved on the host into a register part (a
// mask over the register index: uniform, so the per-pair checks compile to
// scalar branches) and a thread part (a mask over the thread id: one
// predicate per op), so the vector ALU only does the arithmetic.
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

// Moves the tile from layout A to layout B through LDS.  The swizzle is
// linear, so the address of (thread, register) is swz(thread part) ^ lswz[r].
template <class V, int NR, int TLOG>
__device__ __forceinline__ void transpose(V (&v)[NR], V *lds, const DevLayout *A, const DevLayout *B, uint32_t tid)
{
    constexpr int G = sizeof(V) == 16 ? 3 : 4;
    char *base = reinterpret_cast<char *>(lds);
    const uint32_t la = swz_index<G>(thread_local_offset<TLOG>(A, tid)) * (uint32_t)sizeof(V);
#pragma unroll
    for (int r = 0; r < NR; r++)
        *reinterpret_cast<V *>(base + (la ^ A->lswz[r])) = v[r];
    __syncthreads();
    const uint32_t lb = swz_index<G>(thread_local_offset<TLOG>(B, tid)) * (uint32_t)sizeof(V);
#pragma unroll
    for (int r = 0; r < NR; r++)
        v[r] = *reinterpret_cast<const V *>(base + (lb ^ B->lswz[r]));
    __syncthreads();
}

// The pair loops are straight-line code (no branch per pair) so that the
// independent FP64 chains of all pairs interleave; register-bit controls, when
// present, select results instead of branching.
template <int J, class V, class R, int NR>
__device__ __forceinline__ void op_gen(V (&v)[NR], const Mat<R> &M, uint32_t ctrl_reg)
{
    constexpr int bit = 1 << J;
    if (ctrl_reg == 0) {
#pragma unroll
        for (int a = 0; a < NR; a++)
            if (!(a & bit))
                apply2(v[a], v[a | bit], M);
    } else {
#pragma unroll
        for (int a = 0; a < NR; a++) {
            if (a & bit)
                continue;
            V x = v[a], y = v[a | bit];
            apply2(x, y, M);
            const bool ok = ((uint32_t)a & ctrl_reg) == ctrl_reg;
            v[a] = ok ? x : v[a];
            v[a | bit] = ok ? y : v[a | bit];
        }
    }
}

template <int J, class V, int NR> __device__ __forceinline__ void op_swap(V (&v)[NR], uint32_t ctrl_reg)
{
    constexpr int bit = 1 << J;
#pragma unroll
    for (int a = 0; a < NR; a++) {
        if (a & bit)
            continue;
        const bool ok = ((uint32_t)a & ctrl_reg) == ctrl_reg;
        const V x = v[a], y = v[a | bit];
        v[a] = ok ? y : x;
        v[a | bit] = ok ? x : y;
    }
}

template <int J, class V, class R, int NR>
__device__ __forceinline__ void op_diag_reg(V (&v)[NR], R d0r, R d0i, R d1r, R d1i, uint32_t ctrl_reg)
{
    constexpr int bit = 1 << J;
#pragma unroll
    for (int a = 0; a < NR; a++) {
        if (a & bit)
            continue;
        const bool ok = ((uint32_t)a & ctrl_reg) == ctrl_reg;
        const V x = cmul(v[a], d0r, d0i), y = cmul(v[a | bit], d1r, d1i);
        v[a] = ok ? x : v[a];
        v[a | bit] = ok ? y : v[a | bit];
    }
}

// diag(m00, m11) on the op's target, which is a register bit (rbit) or a
// thread bit (tbit).
template <class V, class R, int NR>
__device__ __forceinline__ void op_diag(V (&v)[NR], const DevOp *op, uint32_t tid)
{
    const R d0r = (R)op->m[0], d0i = (R)op->m[1], d1r = (R)op->m[6], d1i = (R)op->m[7];
    const uint32_t ctrl_reg = op->ctrl_reg;
    const uint32_t rb = op->rbit;
    if (rb != kNoReg) {
        switch (rb) {
        case 0: if constexpr (NR > 1) op_diag_reg<0>(v, d0r, d0i, d1r, d1i, ctrl_reg); break;
        case 1: if constexpr (NR > 2) op_diag_reg<1>(v, d0r, d0i, d1r, d1i, ctrl_reg); break;
        case 2: if constexpr (NR > 4) op_diag_reg<2>(v, d0r, d0i, d1r, d1i, ctrl_reg); break;
        case 3: if constexpr (NR > 8) op_diag_reg<3>(v, d0r, d0i, d1r, d1i, ctrl_reg); break;
        case 4: if constexpr (NR > 16) op_diag_reg<4>(v, d0r, d0i, d1r, d1i, ctrl_reg); break;
        default: break;
        }
    } else {
        const bool one = (tid >> op->tbit) & 1u;
        const R dr = one ? d1r : d0r, di = one ? d1i : d0i;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const bool ok = ((uint32_t)r & ctrl_reg) == ctrl_reg;
            const V x = cmul(v[r], dr, di);
            v[r] = ok ? x : v[r];
        }
    }
}

// Applies one GEN or SWAP op whose target is register bit op->rbit.
template <class V, class R, int RLOG>
__device__ __forceinline__ void op_register(V (&v)[1 << RLOG], const DevOp *op)
{
    const uint32_t ctrl_reg = op->ctrl_reg;
    const uint32_t j = op->rbit;
    if (op->kind == OP_GEN) {
        const Mat<R> M = load_mat<R>(op->m);
        switch (j) {
        case 0: if constexpr (RLOG > 0) op_gen<0>(v, M, ctrl_reg); break;
        case 1: if constexpr (RLOG > 1) op_gen<1>(v, M, ctrl_reg); break;
        case 2: if constexpr (RLOG > 2) op_gen<2>(v, M, ctrl_reg); break;
        case 3: if constexpr (RLOG > 3) op_gen<3>(v, M, ctrl_reg); break;
        case 4: if constexpr (RLOG > 4) op_gen<4>(v, M, ctrl_reg); break;
        default: break;
        }
    } else if (op->kind == OP_SWAP) {
        switch (j) {
        case 0: if constexpr (RLOG > 0) op_swap<0>(v, ctrl_reg); break;
        case 1: if constexpr (RLOG > 1) op_swap<1>(v, ctrl_reg); break;
        case 2: if constexpr (RLOG > 2) op_swap<2>(v, ctrl_reg); break;
        case 3: if constexpr (RLOG > 3) op_swap<3>(v, ctrl_reg); break;
        case 4: if constexpr (RLOG > 4) op_swap<4>(v, ctrl_reg); break;
        default: break;
        }
    }
}

__device__ __forceinline__ bool thread_controls_hold(const DevOp *op, uint32_t tid)
{
    const uint32_t ct = op->ctrl_thr;
    return (tid & ct) == ct;
}

template <class V, class R, int RLOG>
__device__ __forceinline__ void run_ops(V (&v)[1 << RLOG], const DevOp *__restrict__ ops, uint32_t ob, uint32_t oe,
                                        uint32_t tid)
{
    for (uint32_t o = ob; o < oe; o++) {
        const DevOp *op = ops + o;
        if (!thread_controls_hold(op, tid))
            continue;
        if (op->kind == OP_DIAG)
            op_diag<V, R>(v, op, tid);
        else
            op_register<V, R, RLOG>(v, op);
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

template <bool NT, class V, int NR>
__device__ __forceinline__ void load_tile(V (&v)[NR], const V *__restrict__ state, uint64_t g, const uint64_t *greg)
{
#pragma unroll
    for (int r = 0; r < NR; r++)
        v[r] = ld<NT>(state + (g | greg[r]));
}

template <bool NT, class V, int NR>
__device__ __forceinline__ void store_tile(const V (&v)[NR], V *__restrict__ state, uint64_t g, const uint64_t *greg)
{
#pragma unroll
    for (int r = 0; r < NR; r++)
        st<NT>(state + (g | greg[r]), v[r]);
}

// Runs the phases of a pass on a loaded tile and leaves it in the store layout.
template <class R, int RLOG, int TLOG>
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
        if (D->transpose_in)
            transpose<V, NR, TLOG>(v, lds, cur, &D->lay, tid);
        cur = &D->lay;
        run_ops<V, R, RLOG>(v, ops, D->op_begin, D->op_end, tid);
    }
    if (P->transpose_st)
        transpose<V, NR, TLOG>(v, lds, cur, &P->lay_st, tid);
}

// One tile of a forward pass: load, run the phases, store.
template <class R, int RLOG, int TLOG, bool NT>
__device__ __forceinline__ void tile_body(typename Cplx<R>::T *__restrict__ state, const DevPass *P,
                                          const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops,
                                          uint64_t tile, typename Cplx<R>::T *lds)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG, K = RLOG + TLOG;
    const uint32_t tid = threadIdx.x;
    const uint64_t g = tile_base<K>(P, tile);
    V v[NR];
    load_tile<NT>(v, state, g | thread_global_offset<TLOG>(P->gthr_ld, tid), P->greg_ld);
    process_tile<R, RLOG, TLOG>(v, P, phases, ops, lds);
    store_tile<NT>(v, state, g | thread_global_offset<TLOG>(P->gthr_st, tid), P->greg_st);
}

// A workgroup's share of a pass, software pipelined: while tile t is
// processed and stored, the loads of the workgroup's next tile are in flight
// (two register buffers, alternating).
template <class R, int RLOG, int TLOG, bool NT>
__device__ __forceinline__ void tile_stream(typename Cplx<R>::T *__restrict__ state, const DevPass *P,
                                            const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops,
                                            uint64_t first, uint64_t stride, uint64_t n_tiles,
                                            typename Cplx<R>::T *lds)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG, K = RLOG + TLOG;
    const uint32_t tid = threadIdx.x;
    const uint64_t tl = thread_global_offset<TLOG>(P->gthr_ld, tid);
    const uint64_t ts = thread_global_offset<TLOG>(P->gthr_st, tid);
    V a[NR], b[NR];
    uint64_t t = first;
    if (t >= n_tiles)
        return;
    load_tile<NT>(a, state, tile_base<K>(P, t) | tl, P->greg_ld);
    for (;;) {
        const uint64_t t2 = t + stride;
        if (t2 < n_tiles)
            load_tile<NT>(b, state, tile_base<K>(P, t2) | tl, P->greg_ld);
        process_tile<R, RLOG, TLOG>(a, P, phases, ops, lds);
        store_tile<NT>(a, state, tile_base<K>(P, t) | ts, P->greg_st);
        if (t2 >= n_tiles)
            break;
        const uint64_t t3 = t2 + stride;
        if (t3 < n_tiles)
            load_tile<NT>(a, state, tile_base<K>(P, t3) | tl, P->greg_ld);
        process_tile<R, RLOG, TLOG>(b, P, phases, ops, lds);
        store_tile<NT>(b, state, tile_base<K>(P, t2) | ts, P->greg_st);
        if (t3 >= n_tiles)
            break;
        t = t3;
    }
}

// --------------------------------------------------------------- adjoint --
// Im<lambda|P|psi> over the thread's register pairs for P = X or Y on bit J.
template <int J, class V, int NR>
__device__ __forceinline__ double ip_pair(const V (&p)[NR], const V (&l)[NR], uint32_t kind)
{
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < NR; a++) {
        if (a & (1 << J))
            continue;
        constexpr int bit = 1 << J;
        const V pa = p[a], pb = p[a | bit], la = l[a], lb = l[a | bit];
        if (kind == OP_IPX)  // conj(la) pb + conj(lb) pa
            acc += (double)la.x * pb.y - (double)la.y * pb.x + (double)lb.x * pa.y - (double)lb.y * pa.x;
        else  // (Y psi)_a = -i psi_b, (Y psi)_b = i psi_a
            acc += -((double)la.x * pb.x + (double)la.y * pb.y) + ((double)lb.x * pa.x + (double)lb.y * pa.y);
    }
    return acc;
}

template <class V, int RLOG>
__device__ __forceinline__ double ip_registers(const V (&p)[1 << RLOG], const V (&l)[1 << RLOG], const DevOp *op,
                                               uint32_t tid)
{
    constexpr int NR = 1 << RLOG;
    double acc = 0.0;
    if (op->kind == OP_IPZ) {
        const uint32_t rb = op->rbit;
        const bool thread_one = rb == kNoReg && ((tid >> op->tbit) & 1u);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const bool one = rb != kNoReg ? (((uint32_t)r >> rb) & 1u) : thread_one;
            const double w = (double)l[r].x * p[r].y - (double)l[r].y * p[r].x;
            acc += one ? -w : w;
        }
        return acc;
    }
    switch (op->rbit) {
    case 0: if constexpr (RLOG > 0) acc = ip_pair<0>(p, l, op->kind); break;
    case 1: if constexpr (RLOG > 1) acc = ip_pair<1>(p, l, op->kind); break;
    case 2: if constexpr (RLOG > 2) acc = ip_pair<2>(p, l, op->kind); break;
    case 3: if constexpr (RLOG > 3) acc = ip_pair<3>(p, l, op->kind); break;
    case 4: if constexpr (RLOG > 4) acc = ip_pair<4>(p, l, op->kind); break;
    default: break;
    }
    return acc;
}

// One tile of a reverse-sweep pass over psi and lambda.  For every IP op the
// workgroup sum is handed to sink(slot, value) in thread 0.
template <class R, int RLOG, int TLOG, class Sink>
__device__ __forceinline__ void tile_body_adjoint(typename Cplx<R>::T *__restrict__ psi,
                                                  typename Cplx<R>::T *__restrict__ lam, const DevPass *P,
                                                  const DevPhase *__restrict__ phases,
                                                  const DevOp *__restrict__ ops, uint64_t tile,
                                                  typename Cplx<R>::T *lds, double *red, Sink sink)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG, K = RLOG + TLOG, NT = 1 << TLOG;
    const uint32_t tid = threadIdx.x;
    const uint64_t g = tile_base<K>(P, tile);
    const uint64_t gl = g | thread_global_offset<TLOG>(P->gthr_ld, tid);
    V p[NR], l[NR];
    load_tile<false>(p, psi, gl, P->greg_ld);
    load_tile<false>(l, lam, gl, P->greg_ld);
    const DevLayout *cur = &P->lay_ld;
    for (uint32_t ph = P->phase_begin; ph < P->phase_end; ph++) {
        const DevPhase *D = phases + ph;
        if (D->transpose_in) {
            transpose<V, NR, TLOG>(p, lds, cur, &D->lay, tid);
            transpose<V, NR, TLOG>(l, lds, cur, &D->lay, tid);
        }
        cur = &D->lay;
        for (uint32_t o = D->op_begin; o < D->op_end; o++) {
            const DevOp *op = ops + o;
            if (op_is_ip(op->kind)) {
                const double s = block_sum<NT>(ip_registers<V, RLOG>(p, l, op, tid), red);
                if (tid == 0)
                    sink(op->slot, s);
                continue;
            }
            if (!thread_controls_hold(op, tid))
                continue;
            if (op->kind == OP_DIAG) {
                op_diag<V, R>(p, op, tid);
                op_diag<V, R>(l, op, tid);
            } else {
                op_register<V, R, RLOG>(p, op);
                op_register<V, R, RLOG>(l, op);
            }
        }
    }
    if (P->transpose_st) {
        transpose<V, NR, TLOG>(p, lds, cur, &P->lay_st, tid);
        transpose<V, NR, TLOG>(l, lds, cur, &P->lay_st, tid);
    }
    const uint64_t gs = g | thread_global_offset<TLOG>(P->gthr_st, tid);
    store_tile<false>(p, psi, gs, P->greg_st);
    store_tile<false>(l, lam, gs, P->greg_st);
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

// This thread's part of Im<lambda|P|psi> for an IP op on a whole LDS state.
template <class V>
__device__ __forceinline__ double small_ip(const V *sp, const V *sl, uint32_t n, const DevOp *op)
{
    const uint32_t n_amp = 1u << n, t = op->target;
    double acc = 0.0;
    if (op->kind == OP_IPZ) {
        for (uint32_t i = threadIdx.x; i < n_amp; i += blockDim.x) {
            const double sg = (i >> t) & 1u ? -1.0 : 1.0;
            acc += sg * ((double)sl[i].x * sp[i].y - (double)sl[i].y * sp[i].x);
        }
    } else {
        for (uint32_t p = threadIdx.x; p < n_amp / 2; p += blockDim.x) {
            const uint32_t i0 = (uint32_t)insert_zero_bit(p, t), i1 = i0 | (1u << t);
            const V pa = sp[i0], pb = sp[i1], la = sl[i0], lb = sl[i1];
            if (op->kind == OP_IPX)
                acc += (double)la.x * pb.y - (double)la.y * pb.x + (double)lb.x * pa.y - (double)lb.y * pa.x;
            else
                acc += -((double)la.x * pb.x + (double)la.y * pb.y) + ((double)lb.x * pa.x + (double)lb.y * pa.y);
        }
    }
    return acc;
}

}  // namespace dmv
