// Device helpers shared by the HIP kernels: complex amplitude types, loads and
// stores, 2x2 updates and wave64/workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace dmv {

using c128 = double __attribute__((ext_vector_type(2)));
using c64 = float __attribute__((ext_vector_type(2)));

template <class R> struct Cplx;
template <> struct Cplx<double> { using T = c128; };
template <> struct Cplx<float> { using T = c64; };

template <bool NT, class V> __device__ __forceinline__ V ld(const V *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool NT, class V> __device__ __forceinline__ void st(V *p, V v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// The 2x2 matrix of an op in the precision of the state.
template <class R> struct Mat {
    R m00r, m00i, m01r, m01i, m10r, m10i, m11r, m11i;
};

template <class R> __host__ __device__ __forceinline__ Mat<R> load_mat(const double *m)
{
    Mat<R> M;
    M.m00r = (R)m[0]; M.m00i = (R)m[1]; M.m01r = (R)m[2]; M.m01i = (R)m[3];
    M.m10r = (R)m[4]; M.m10i = (R)m[5]; M.m11r = (R)m[6]; M.m11i = (R)m[7];
    return M;
}

// (a, b) <- M (a, b)
template <class V, class R> __device__ __forceinline__ void apply2(V &a, V &b, const Mat<R> &M)
{
    V na, nb;
    na.x = M.m00r * a.x - M.m00i * a.y + M.m01r * b.x - M.m01i * b.y;
    na.y = M.m00r * a.y + M.m00i * a.x + M.m01r * b.y + M.m01i * b.x;
    nb.x = M.m10r * a.x - M.m10i * a.y + M.m11r * b.x - M.m11i * b.y;
    nb.y = M.m10r * a.y + M.m10i * a.x + M.m11r * b.y + M.m11i * b.x;
    a = na;
    b = nb;
}

template <class V, class R> __device__ __forceinline__ V cmul(V a, R dr, R di)
{
    V r;
    r.x = a.x * dr - a.y * di;
    r.y = a.x * di + a.y * dr;
    return r;
}

template <class V> __device__ __forceinline__ double norm2(V a)
{
    return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// Sum over the 64 lanes of a wave (every lane gets the result).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// The same without LDS round trips: DPP moves inside each row of 16 lanes
// (xor 1 and 2 by quad permutes, then the half-row and row mirrors, each
// bringing the other half's sum once the halves are uniform), then the
// gfx950 row swaps: permlane16_swap of a value with itself returns rows
// (0,0,2,2) and (1,1,3,3), whose sum is the pair sum; permlane32_swap
// returns halves (lo,lo) and (hi,hi).  Every lane ends with the same value
// (each step adds the same two terms in every lane).
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v)
{
    int2 x = __builtin_bit_cast(int2, v);
    x.x = __builtin_amdgcn_mov_dpp(x.x, CTRL, 0xF, 0xF, false);
    x.y = __builtin_amdgcn_mov_dpp(x.y, CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, x);
}

__device__ __forceinline__ double wave_sum_rows(double v)
{
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    v += dpp_mov<0x140>(v);  // row_mirror
    int2 x = __builtin_bit_cast(int2, v);
    const auto a = __builtin_amdgcn_permlane16_swap(x.x, x.x, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(x.y, x.y, false, false);
    v = __builtin_bit_cast(double, int2{(int)a[0], (int)b[0]}) + __builtin_bit_cast(double, int2{(int)a[1], (int)b[1]});
    x = __builtin_bit_cast(int2, v);
    const auto c = __builtin_amdgcn_permlane32_swap(x.x, x.x, false, false);
    const auto d = __builtin_amdgcn_permlane32_swap(x.y, x.y, false, false);
    return __builtin_bit_cast(double, int2{(int)c[0], (int)d[0]}) + __builtin_bit_cast(double, int2{(int)c[1], (int)d[1]});
}

// Sum over a workgroup of NT threads; `scratch` holds NT/64 doubles.  The
// result is valid in every thread.  Fixed order, so deterministic.
template <int NT> __device__ __forceinline__ double block_sum(double v, double *scratch)
{
    v = wave_sum(v);
    constexpr int NW = NT / 64;
    if constexpr (NW == 1) {
        return v;
    } else {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        __syncthreads();
        if (lane == 0)
            scratch[w] = v;
        __syncthreads();
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NW; i++)
            s += scratch[i];
        return s;
    }
}

// Host-visible launch helpers (defined in runtime.hip).
hipStream_t stream();
void prof_begin(int kernel_class);
void prof_end(int kernel_class, double bytes);
int check_launch(const char *what);
int64_t tune_value(const char *name, int64_t dflt);
int gate_pairs_per_thread();
bool gate_nontemporal(uint32_t target, int dtype);
int device_cus();  // compute units of the current device

}  // namespace dmv
