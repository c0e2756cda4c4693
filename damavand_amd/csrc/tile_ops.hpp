// Device building blocks of the fused tile executor, shared by the big-state
// pass kernels, the whole-state kernels and the batched persistent kernel.
#pragma once

#include "device.hpp"

namespace dmv {

// LDS address swizzle: folds the higher G-bit groups of the local index into
// its low G bits.  A local bit p lands on swizzled bit p mod G, so lanes that
// differ in bits of distinct residues never share a 128-bit bank group.
template <int G> __device__ __forceinline__ uint32_t swz(uint32_t li)
{
    uint32_t f = 0;
#pragma unroll
    for (int s = G; s < 16; s += G)
        f ^= li >> s;
    return li ^ (f & ((1u << G) - 1));
}

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

template <class V, int NR, int TLOG, int G>
__device__ __forceinline__ void transpose(V (&v)[NR], V *lds, const DevLayout *A, const DevLayout *B, uint32_t tid)
{
    const uint32_t la = thread_local_offset<TLOG>(A, tid);
#pragma unroll
    for (int r = 0; r < NR; r++)
        lds[swz<G>(la | A->lreg[r])] = v[r];
    __syncthreads();
    const uint32_t lb = thread_local_offset<TLOG>(B, tid);
#pragma unroll
    for (int r = 0; r < NR; r++)
        v[r] = lds[swz<G>(lb | B->lreg[r])];
    __syncthreads();
}

template <int J, class V, class R, int NR>
__device__ __forceinline__ void op_gen(V (&v)[NR], const Mat<R> &M, uint32_t ctrl, uint32_t lt,
                                       const uint32_t (&lreg)[NR])
{
#pragma unroll
    for (int a = 0; a < NR; a++) {
        if (a & (1 << J))
            continue;
        constexpr int bit = 1 << J;
        if (ctrl == 0 || ((lt | lreg[a]) & ctrl) == ctrl)
            apply2(v[a], v[a | bit], M);
    }
}

template <int J, class V, int NR>
__device__ __forceinline__ void op_swap(V (&v)[NR], uint32_t ctrl, uint32_t lt, const uint32_t (&lreg)[NR])
{
#pragma unroll
    for (int a = 0; a < NR; a++) {
        if (a & (1 << J))
            continue;
        constexpr int bit = 1 << J;
        if (ctrl == 0 || ((lt | lreg[a]) & ctrl) == ctrl) {
            V t = v[a];
            v[a] = v[a | bit];
            v[a | bit] = t;
        }
    }
}

template <class V, class R, int NR>
__device__ __forceinline__ void op_diag(V (&v)[NR], const DevOp *op, uint32_t lt, const uint32_t (&lreg)[NR])
{
    const R d0r = (R)op->m[0], d0i = (R)op->m[1], d1r = (R)op->m[6], d1i = (R)op->m[7];
    const uint32_t ctrl = (uint32_t)op->ctrl, t = op->target;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const uint32_t li = lt | lreg[r];
        if ((li & ctrl) == ctrl) {
            const bool one = (li >> t) & 1u;
            v[r] = cmul(v[r], one ? d1r : d0r, one ? d1i : d0i);
        }
    }
}

// Applies one GEN or SWAP op whose target is register bit op->rbit.
template <class V, class R, int RLOG>
__device__ __forceinline__ void op_register(V (&v)[1 << RLOG], const DevOp *op, uint32_t lt,
                                            const uint32_t (&lreg)[1 << RLOG])
{
    const uint32_t ctrl = (uint32_t)op->ctrl;
    const uint32_t j = op->rbit;
    if (op->kind == OP_GEN) {
        const Mat<R> M = load_mat<R>(op->m);
        switch (j) {
        case 0: if constexpr (RLOG > 0) op_gen<0>(v, M, ctrl, lt, lreg); break;
        case 1: if constexpr (RLOG > 1) op_gen<1>(v, M, ctrl, lt, lreg); break;
        case 2: if constexpr (RLOG > 2) op_gen<2>(v, M, ctrl, lt, lreg); break;
        case 3: if constexpr (RLOG > 3) op_gen<3>(v, M, ctrl, lt, lreg); break;
        case 4: if constexpr (RLOG > 4) op_gen<4>(v, M, ctrl, lt, lreg); break;
        default: break;
        }
    } else if (op->kind == OP_SWAP) {
        switch (j) {
        case 0: if constexpr (RLOG > 0) op_swap<0>(v, ctrl, lt, lreg); break;
        case 1: if constexpr (RLOG > 1) op_swap<1>(v, ctrl, lt, lreg); break;
        case 2: if constexpr (RLOG > 2) op_swap<2>(v, ctrl, lt, lreg); break;
        case 3: if constexpr (RLOG > 3) op_swap<3>(v, ctrl, lt, lreg); break;
        case 4: if constexpr (RLOG > 4) op_swap<4>(v, ctrl, lt, lreg); break;
        default: break;
        }
    }
}

template <class V, class R, int RLOG>
__device__ __forceinline__ void run_ops(V (&v)[1 << RLOG], const DevOp *__restrict__ ops, uint32_t ob, uint32_t oe,
                                        uint32_t lt, const uint32_t (&lreg)[1 << RLOG])
{
    for (uint32_t o = ob; o < oe; o++) {
        const DevOp *op = ops + o;
        if (op->kind == OP_DIAG)
            op_diag<V, R>(v, op, lt, lreg);
        else
            op_register<V, R, RLOG>(v, op, lt, lreg);
    }
}

// One tile of a forward pass: load, run the phases, store.
template <class R, int RLOG, int TLOG, bool NT>
__device__ __forceinline__ void tile_body(typename Cplx<R>::T *__restrict__ state, const DevPass *P,
                                          const DevPhase *__restrict__ phases, const DevOp *__restrict__ ops,
                                          uint64_t tile, typename Cplx<R>::T *lds)
{
    using V = typename Cplx<R>::T;
    constexpr int NR = 1 << RLOG, K = RLOG + TLOG;
    constexpr int G = sizeof(V) == 16 ? 3 : 4;
    const uint32_t tid = threadIdx.x;
    uint64_t g = tile;
#pragma unroll
    for (int j = 0; j < K; j++)
        g = insert_zero_bit(g, P->qpos[j]);
    const uint64_t gl = g | thread_global_offset<TLOG>(P->gthr_ld, tid);
    V v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++)
        v[r] = ld<NT>(state + (gl | P->greg_ld[r]));
    const DevLayout *cur = &P->lay_ld;
    for (uint32_t ph = P->phase_begin; ph < P->phase_end; ph++) {
        const DevPhase *D = phases + ph;
        if (D->transpose_in)
            transpose<V, NR, TLOG, G>(v, lds, cur, &D->lay, tid);
        cur = &D->lay;
        uint32_t lreg[NR];
#pragma unroll
        for (int r = 0; r < NR; r++)
            lreg[r] = cur->lreg[r];
        run_ops<V, R, RLOG>(v, ops, D->op_begin, D->op_end, thread_local_offset<TLOG>(cur, tid), lreg);
    }
    if (P->transpose_st)
        transpose<V, NR, TLOG, G>(v, lds, cur, &P->lay_st, tid);
    const uint64_t gs = g | thread_global_offset<TLOG>(P->gthr_st, tid);
#pragma unroll
    for (int r = 0; r < NR; r++)
        st<NT>(state + (gs | P->greg_st[r]), v[r]);
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
                                               uint32_t lt, const uint32_t (&lreg)[1 << RLOG])
{
    constexpr int NR = 1 << RLOG;
    double acc = 0.0;
    if (op->kind == OP_IPZ) {
        const uint32_t t = op->target;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const double s = ((lt | lreg[r]) >> t) & 1u ? -1.0 : 1.0;
            acc += s * ((double)l[r].x * p[r].y - (double)l[r].y * p[r].x);
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
    constexpr int G = sizeof(V) == 16 ? 3 : 4;
    const uint32_t tid = threadIdx.x;
    uint64_t g = tile;
#pragma unroll
    for (int j = 0; j < K; j++)
        g = insert_zero_bit(g, P->qpos[j]);
    const uint64_t gl = g | thread_global_offset<TLOG>(P->gthr_ld, tid);
    V p[NR], l[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        p[r] = psi[gl | P->greg_ld[r]];
        l[r] = lam[gl | P->greg_ld[r]];
    }
    const DevLayout *cur = &P->lay_ld;
    for (uint32_t ph = P->phase_begin; ph < P->phase_end; ph++) {
        const DevPhase *D = phases + ph;
        if (D->transpose_in) {
            transpose<V, NR, TLOG, G>(p, lds, cur, &D->lay, tid);
            transpose<V, NR, TLOG, G>(l, lds, cur, &D->lay, tid);
        }
        cur = &D->lay;
        uint32_t lreg[NR];
#pragma unroll
        for (int r = 0; r < NR; r++)
            lreg[r] = cur->lreg[r];
        const uint32_t lt = thread_local_offset<TLOG>(cur, tid);
        for (uint32_t o = D->op_begin; o < D->op_end; o++) {
            const DevOp *op = ops + o;
            if (op_is_ip(op->kind)) {
                const double s = block_sum<NT>(ip_registers<V, RLOG>(p, l, op, lt, lreg), red);
                if (tid == 0)
                    sink(op->slot, s);
            } else if (op->kind == OP_DIAG) {
                op_diag<V, R>(p, op, lt, lreg);
                op_diag<V, R>(l, op, lt, lreg);
            } else {
                op_register<V, R, RLOG>(p, op, lt, lreg);
                op_register<V, R, RLOG>(l, op, lt, lreg);
            }
        }
    }
    if (P->transpose_st) {
        transpose<V, NR, TLOG, G>(p, lds, cur, &P->lay_st, tid);
        transpose<V, NR, TLOG, G>(l, lds, cur, &P->lay_st, tid);
    }
    const uint64_t gs = g | thread_global_offset<TLOG>(P->gthr_st, tid);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        psi[gs | P->greg_st[r]] = p[r];
        lam[gs | P->greg_st[r]] = l[r];
    }
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
