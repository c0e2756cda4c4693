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
