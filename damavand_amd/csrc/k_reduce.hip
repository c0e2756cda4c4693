// State utilities, reductions and the sampler.  Every reduction sums in a
// fixed order (per-thread strided partials, wave64 shuffle tree, per-block
// partials, then one deterministic workgroup), so results repeat bit for bit.
#include "device.hpp"
#include "kernels.hpp"

namespace dmv {

constexpr int RT = 256;          // threads of a reduction block
constexpr int kBlockBits = 12;   // amplitudes per reduction block: 4096

static uint32_t grid_for(uint64_t items, uint64_t per_block, uint32_t cap = 1u << 20)
{
    uint64_t g = (items + per_block - 1) / per_block;
    if (g < 1)
        g = 1;
    return (uint32_t)(g > cap ? cap : g);
}

// ------------------------------------------------------------- elementwise --
template <class V>
__global__ void k_init_basis(V *__restrict__ s, uint64_t n_amp, uint64_t basis, StateLayout L)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_amp;
         i += (uint64_t)gridDim.x * blockDim.x) {
        V v;
        v.x = i == basis ? 1 : 0;
        v.y = 0;
        s[elem(i, L)] = v;
    }
}

int launch_init_basis(int dtype, void *state, int n, StateLayout L, uint64_t basis)
{
    const uint64_t n_amp = uint64_t(1) << n;
    DMV_CHECK_ARG(basis < n_amp || basis == ~uint64_t(0), "basis index out of range");
    const uint32_t g = grid_for(n_amp, RT * 8);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_init_basis<c64>, dim3(g), dim3(RT), 0, stream(), (c64 *)state, n_amp, basis, L);
    else
        hipLaunchKernelGGL(k_init_basis<c128>, dim3(g), dim3(RT), 0, stream(), (c128 *)state, n_amp, basis, L);
    return check_launch("init");
}

// Each workgroup copies one block of 4 x 256 16-byte elements, four loads in
// flight per thread, streaming (nontemporal) both ways: the fastest of the
// copy shapes measured on 16 GiB buffers (5.67 TB/s against 5.33 TB/s for a
// grid-stride loop and 5.27 TB/s for 16 elements per thread).
constexpr int kCopyUnroll = 4;

__global__ void __launch_bounds__(RT) k_copy(c128 *__restrict__ d, const c128 *__restrict__ s, uint64_t n)
{
    const uint64_t base = (uint64_t)blockIdx.x * (RT * kCopyUnroll) + threadIdx.x;
    c128 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++)
        if (base + (uint64_t)u * RT < n)
            v[u] = __builtin_nontemporal_load(s + base + (uint64_t)u * RT);
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++)
        if (base + (uint64_t)u * RT < n)
            __builtin_nontemporal_store(v[u], d + base + (uint64_t)u * RT);
}

__global__ void k_copy_bytes(uint8_t *__restrict__ d, const uint8_t *__restrict__ s, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

int launch_copy(void *dst, const void *src, uint64_t bytes)
{
    const uint64_t n16 = bytes / 16;
    if (n16) {
        const uint64_t blocks = (n16 + RT * kCopyUnroll - 1) / (RT * kCopyUnroll);
        DMV_CHECK_ARG(blocks <= 0xFFFFFFFFull, "copy too large");
        hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(RT), 0, stream(), (c128 *)dst, (const c128 *)src,
                           n16);
    }
    const uint64_t rest = bytes - n16 * 16;
    if (rest)
        hipLaunchKernelGGL(k_copy_bytes, dim3(1), dim3(64), 0, stream(), (uint8_t *)dst + n16 * 16,
                           (const uint8_t *)src + n16 * 16, rest);
    return check_launch("copy");
}

template <class V>
__global__ void k_split(const V *__restrict__ s, uint64_t begin, uint64_t count, double *re, double *im,
                        StateLayout L)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const V v = s[elem(begin + i, L)];
        re[i] = v.x;
        im[i] = v.y;
    }
}

template <class V>
__global__ void k_join(V *__restrict__ s, uint64_t begin, uint64_t count, const double *re, const double *im,
                       StateLayout L)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        V v;
        v.x = re[i];
        v.y = im[i];
        s[elem(begin + i, L)] = v;
    }
}

template <class V>
__global__ void k_probs(const V *__restrict__ s, uint64_t begin, uint64_t count, double *out, StateLayout L)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x) {
        const V v = s[elem(begin + i, L)];
        out[i] = (double)v.x * v.x + (double)v.y * v.y;
    }
}

int launch_split(int dtype, const void *state, StateLayout L, uint64_t begin, uint64_t count, double *re, double *im)
{
    const uint32_t g = grid_for(count, RT * 4);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_split<c64>, dim3(g), dim3(RT), 0, stream(), (const c64 *)state, begin, count, re, im, L);
    else
        hipLaunchKernelGGL(k_split<c128>, dim3(g), dim3(RT), 0, stream(), (const c128 *)state, begin, count, re, im, L);
    return check_launch("split");
}

int launch_join(int dtype, void *state, StateLayout L, uint64_t begin, uint64_t count, const double *re,
                const double *im)
{
    const uint32_t g = grid_for(count, RT * 4);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_join<c64>, dim3(g), dim3(RT), 0, stream(), (c64 *)state, begin, count, re, im, L);
    else
        hipLaunchKernelGGL(k_join<c128>, dim3(g), dim3(RT), 0, stream(), (c128 *)state, begin, count, re, im, L);
    return check_launch("join");
}

int launch_probs(int dtype, const void *state, StateLayout L, uint64_t begin, uint64_t count, double *out)
{
    const uint32_t g = grid_for(count, RT * 4);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_probs<c64>, dim3(g), dim3(RT), 0, stream(), (const c64 *)state, begin, count, out, L);
    else
        hipLaunchKernelGGL(k_probs<c128>, dim3(g), dim3(RT), 0, stream(), (const c128 *)state, begin, count, out, L);
    return check_launch("probabilities");
}

// --------------------------------------------------------------- reductions --
template <class V>
__global__ void __launch_bounds__(RT) k_block_norms(const V *__restrict__ s, uint32_t bb, double *partials,
                                                   StateLayout L)
{
    __shared__ double red[RT / 64];
    const uint64_t eb = elem((uint64_t)blockIdx.x << bb, L);  // a block lies in one span: element eb ^ i
    const uint32_t span = 1u << bb;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < span; i += RT)
        acc += norm2(s[eb ^ i]);
    acc = block_sum<RT>(acc, red);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = acc;
}

int launch_block_norms(int dtype, const void *state, int n, StateLayout L, int bb, double *partials)
{
    DMV_CHECK_ARG(bb >= 0 && bb <= n && n - bb <= 31 && (L.shift == 0 || bb <= (int)L.shift), "block bits out of range");
    const uint32_t blocks = 1u << (n - bb);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_block_norms<c64>, dim3(blocks), dim3(RT), 0, stream(), (const c64 *)state, (uint32_t)bb,
                           partials, L);
    else
        hipLaunchKernelGGL(k_block_norms<c128>, dim3(blocks), dim3(RT), 0, stream(), (const c128 *)state,
                           (uint32_t)bb, partials, L);
    return check_launch("block norms");
}

// out[v] = sum_i in[i * stride + v] for v < nv (nv <= 4), one workgroup.
__global__ void __launch_bounds__(RT) k_sum(const double *__restrict__ in, uint64_t count, uint32_t stride,
                                            uint32_t nv, double *out)
{
    __shared__ double red[RT / 64];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t i = threadIdx.x; i < count; i += RT) {
#pragma unroll
        for (int v = 0; v < 4; v++)
            if ((uint32_t)v < nv)
                acc[v] += in[i * stride + v];
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const double s = block_sum<RT>(acc[v], red);
        if (threadIdx.x == 0 && (uint32_t)v < nv)
            out[v] = s;
    }
}

int launch_sum(const double *in, uint64_t count, double *out)
{
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(RT), 0, stream(), in, count, 1u, 1u, out);
    return check_launch("sum");
}

// <Z_q>: each workgroup reads 2^G consecutive blocks of 4096 amplitudes
// (G = group bits) and writes the total mass of its 2^(12+G) amplitudes and,
// for every q < 12 + G, the mass with bit q set; bits >= 12 + G are constant
// over a workgroup and are added by k_expect_z_final.  Thread t reads
// amplitudes t + 256 k (k < 16) of every block, so bits 0-7 of an index are
// the thread's (bits 0-5 the lane, 6-7 the wave) and bits 8-11 are k's: a
// thread keeps only its total, the four k-bit masses and the G group-bit
// masses, and the lane-bit masses come from the wave totals.  One barrier
// per workgroup.
constexpr int kZGroupMax = 4;
constexpr int kZStrideMax = 1 + kBlockBits + kZGroupMax;  // doubles per workgroup partial
static_assert(RT == 256 && RT * 16 == (1 << kBlockBits), "expect_z thread mapping");

template <class V>
__global__ void __launch_bounds__(RT) k_expect_z_part(const V *__restrict__ s, uint32_t B, uint32_t G,
                                                     double *partials, StateLayout L)
{
    constexpr int NW = RT / 64;
    constexpr int NV = 1 + 6 + 4 + kZGroupMax;  // total, lane bits, k bits, group bits
    __shared__ double red[NW][NV];
    const uint32_t bw = B + G;
    const uint64_t eb = elem((uint64_t)blockIdx.x << bw, L);  // the group lies in one span
    const uint32_t span = 1u << B;
    double tot = 0.0, hk[4] = {0.0, 0.0, 0.0, 0.0}, hg[kZGroupMax] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t g = 0; g < (1u << G); g++) {
        double p[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const uint32_t i = threadIdx.x + (uint32_t)k * RT;
            p[k] = i < span ? norm2(s[eb ^ ((uint64_t)g << B) ^ i]) : 0.0;
        }
        double bt = 0.0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            bt += p[k];
#pragma unroll
            for (int j = 0; j < 4; j++)
                if ((k >> j) & 1)
                    hk[j] += p[k];
        }
        tot += bt;
#pragma unroll
        for (int j = 0; j < kZGroupMax; j++)
            if ((g >> j) & 1u)
                hg[j] += bt;
    }
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double v[NV];
    v[0] = wave_sum(tot);
#pragma unroll
    for (int q = 0; q < 6; q++)
        v[1 + q] = wave_sum(((lane >> q) & 1u) ? tot : 0.0);
#pragma unroll
    for (int j = 0; j < 4; j++)
        v[7 + j] = wave_sum(hk[j]);
#pragma unroll
    for (int j = 0; j < kZGroupMax; j++)
        v[11 + j] = wave_sum(hg[j]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++)
            red[w][i] = v[i];
    }
    __syncthreads();
    // thread 0: total; thread 1 + q: mass with bit q set (q < B + G)
    const int q = (int)threadIdx.x - 1;
    if (q < (int)bw) {
        double acc = 0.0;
#pragma unroll
        for (int ww = 0; ww < NW; ww++) {
            double x;
            if (q < 0)
                x = red[ww][0];
            else if (q < 6)
                x = red[ww][1 + q];
            else if (q < 8)
                x = ((ww >> (q - 6)) & 1) ? red[ww][0] : 0.0;
            else if (q < kBlockBits)
                x = red[ww][7 + q - 8];
            else
                x = red[ww][11 + q - kBlockBits];
            acc += x;
        }
        partials[(uint64_t)blockIdx.x * (1 + bw) + threadIdx.x] = acc;
    }
}

// Sums the workgroup partials (stride 1 + bw, bw = bits inside a workgroup)
// into <Z_q> = total - 2 mass(bit q set), one workgroup, fixed order.
__global__ void __launch_bounds__(RT) k_expect_z_final(const double *__restrict__ partials, uint64_t ngroups, uint32_t bw,
                                                       uint32_t n, double *out)
{
    __shared__ double red[RT / 64];
    double tot = 0.0, m[kMaxQubits];
#pragma unroll
    for (int q = 0; q < kMaxQubits; q++)
        m[q] = 0.0;
    for (uint64_t b = threadIdx.x; b < ngroups; b += RT) {
        const double *pb = partials + b * (1 + bw);
        const double t = pb[0];
        tot += t;
#pragma unroll
        for (int q = 0; q < kMaxQubits; q++) {
            if ((uint32_t)q < bw)
                m[q] += pb[1 + (q < kZStrideMax - 1 ? q : 0)];
            else if ((uint32_t)q < n && ((b >> (q - bw)) & 1u))
                m[q] += t;
        }
    }
    const double T = block_sum<RT>(tot, red);
#pragma unroll
    for (int q = 0; q < kMaxQubits; q++) {
        if ((uint32_t)q >= n)
            break;
        const double v = block_sum<RT>(m[q], red);
        if (threadIdx.x == 0)
            out[q] = T - 2.0 * v;
    }
}

// Group bits of <Z_q>: up to four blocks per workgroup while the grid keeps
// at least 2^11 workgroups, within one span of the layout.
static uint32_t expect_z_group_bits(int n, int B, StateLayout L)
{
    int g = n - B - 11;
    g = g < 0 ? 0 : (g > kZGroupMax ? kZGroupMax : g);
    if (L.shift && B + g > (int)L.shift)
        g = (int)L.shift - B;
    return (uint32_t)g;
}

int launch_expect_z(int dtype, const void *state, int n, StateLayout L, double *partials, double *out)
{
    const int B = n < kBlockBits ? n : kBlockBits;
    const uint32_t G = expect_z_group_bits(n, B, L);
    const uint64_t ngroups = uint64_t(1) << (n - B - (int)G);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_expect_z_part<c64>, dim3((uint32_t)ngroups), dim3(RT), 0, stream(), (const c64 *)state,
                           (uint32_t)B, G, partials, L);
    else
        hipLaunchKernelGGL(k_expect_z_part<c128>, dim3((uint32_t)ngroups), dim3(RT), 0, stream(), (const c128 *)state,
                           (uint32_t)B, G, partials, L);
    DMV_TRY(check_launch("<Z> partials"));
    hipLaunchKernelGGL(k_expect_z_final, dim3(1), dim3(RT), 0, stream(), partials, ngroups, (uint32_t)B + G, (uint32_t)n,
                       out);
    return check_launch("<Z> final");
}

// Z/ZZ observable h(x) = sum_q a_q z_q + sum_t b_t z_{i_t} z_{j_t}, z = 1 - 2 bit.
struct ObsTerms {
    const double *a;
    const int *pi;
    const int *pj;
    const double *b;
    int n_pairs;
};

constexpr int kMaxPairs = 128;

template <class V, bool WRITE>
__global__ void __launch_bounds__(RT) k_obs(const V *__restrict__ s, V *__restrict__ lam, uint32_t n,
                                            uint32_t B, ObsTerms T, double *partials, StateLayout L)
{
    __shared__ double red[RT / 64];
    __shared__ double sa[kMaxQubits];
    __shared__ double sb[kMaxPairs];
    __shared__ uint8_t spi[kMaxPairs], spj[kMaxPairs];
    for (uint32_t q = threadIdx.x; q < n; q += RT)
        sa[q] = T.a ? T.a[q] : 0.0;
    for (int t = threadIdx.x; t < T.n_pairs; t += RT) {
        sb[t] = T.b[t];
        spi[t] = (uint8_t)T.pi[t];
        spj[t] = (uint8_t)T.pj[t];
    }
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x << B;
    const uint64_t e0 = elem(base, L);
    const uint32_t span = 1u << B;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < span; i += RT) {
        const uint64_t x = base + i;
        double h = 0.0;
        for (uint32_t q = 0; q < n; q++)
            h += ((x >> q) & 1u) ? -sa[q] : sa[q];
        for (int t = 0; t < T.n_pairs; t++)
            h += (((x >> spi[t]) ^ (x >> spj[t])) & 1u) ? -sb[t] : sb[t];
        const V v = s[e0 ^ i];
        acc += h * norm2(v);
        if constexpr (WRITE) {
            V w;
            w.x = v.x * h;
            w.y = v.y * h;
            lam[e0 ^ i] = w;
        }
    }
    acc = block_sum<RT>(acc, red);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = acc;
}

template <bool WRITE>
static int run_obs(int dtype, const void *psi, void *lam, int n, StateLayout L, const double *a, int n_pairs, const int *pi,
                   const int *pj, const double *b, double *partials, double *out)
{
    DMV_CHECK_ARG(n_pairs >= 0 && n_pairs <= kMaxPairs, "at most %d ZZ terms", kMaxPairs);
    const int B = n < kBlockBits ? n : kBlockBits;
    const uint64_t nblk = uint64_t(1) << (n - B);
    ObsTerms T{a, pi, pj, b, n_pairs};
    if (dtype == DMV_C64)
        hipLaunchKernelGGL((k_obs<c64, WRITE>), dim3((uint32_t)nblk), dim3(RT), 0, stream(), (const c64 *)psi,
                           (c64 *)lam, (uint32_t)n, (uint32_t)B, T, partials, L);
    else
        hipLaunchKernelGGL((k_obs<c128, WRITE>), dim3((uint32_t)nblk), dim3(RT), 0, stream(), (const c128 *)psi,
                           (c128 *)lam, (uint32_t)n, (uint32_t)B, T, partials, L);
    DMV_TRY(check_launch("<H> partials"));
    return launch_sum(partials, nblk, out);
}

int launch_expect_zz(int dtype, const void *state, int n, StateLayout L, const double *a, int n_pairs, const int *pi,
                     const int *pj, const double *b, double *partials, double *out)
{
    return run_obs<false>(dtype, state, nullptr, n, L, a, n_pairs, pi, pj, b, partials, out);
}

int launch_apply_obs(int dtype, const void *psi, void *lam, int n, StateLayout L, const double *a, int n_pairs, const int *pi,
                     const int *pj, const double *b, double *partials, double *out)
{
    return run_obs<true>(dtype, psi, lam, n, L, a, n_pairs, pi, pj, b, partials, out);
}

// Diagonal observable given by a table over up to kMaxTableWires qubits:
// h(x) = table[sum_j bit_{wires[j]}(x) 2^j].  <H> partials per block, and
// with WRITE also lambda = h * psi.
template <class V, bool WRITE>
__global__ void __launch_bounds__(RT) k_obs_table(const V *__restrict__ s, V *__restrict__ lam, uint32_t B,
                                                  const double *__restrict__ table, TableWires W, double *partials,
                                                  StateLayout L)
{
    __shared__ double red[RT / 64];
    __shared__ double st[1 << kMaxTableWires];
    for (uint32_t i = threadIdx.x; i < (1u << W.n); i += RT)
        st[i] = table[i];
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x << B;
    const uint64_t e0 = elem(base, L);
    const uint32_t span = 1u << B;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < span; i += RT) {
        const uint64_t x = base + i;
        uint32_t idx = 0;
        for (uint32_t j = 0; j < W.n; j++)
            idx |= (uint32_t)((x >> W.q[j]) & 1u) << j;
        const double h = st[idx];
        const V v = s[e0 ^ i];
        acc += h * norm2(v);
        if constexpr (WRITE) {
            V w;
            w.x = v.x * h;
            w.y = v.y * h;
            lam[e0 ^ i] = w;
        }
    }
    acc = block_sum<RT>(acc, red);
    if (threadIdx.x == 0)
        partials[blockIdx.x] = acc;
}

int launch_obs_table(int dtype, const void *psi, void *lam, int n, StateLayout L, const double *table,
                     const TableWires &W, double *partials, double *out)
{
    DMV_CHECK_ARG(W.n <= (uint32_t)kMaxTableWires, "at most %d wires in a table observable", kMaxTableWires);
    for (uint32_t j = 0; j < W.n; j++)
        DMV_CHECK_ARG(W.q[j] < (uint32_t)n, "table observable wire out of range");
    const int B = n < kBlockBits ? n : kBlockBits;
    const uint64_t nblk = uint64_t(1) << (n - B);
    if (dtype == DMV_C64) {
        if (lam)
            hipLaunchKernelGGL((k_obs_table<c64, true>), dim3((uint32_t)nblk), dim3(RT), 0, stream(), (const c64 *)psi,
                               (c64 *)lam, (uint32_t)B, table, W, partials, L);
        else
            hipLaunchKernelGGL((k_obs_table<c64, false>), dim3((uint32_t)nblk), dim3(RT), 0, stream(),
                               (const c64 *)psi, (c64 *)nullptr, (uint32_t)B, table, W, partials, L);
    } else {
        if (lam)
            hipLaunchKernelGGL((k_obs_table<c128, true>), dim3((uint32_t)nblk), dim3(RT), 0, stream(),
                               (const c128 *)psi, (c128 *)lam, (uint32_t)B, table, W, partials, L);
        else
            hipLaunchKernelGGL((k_obs_table<c128, false>), dim3((uint32_t)nblk), dim3(RT), 0, stream(),
                               (const c128 *)psi, (c128 *)nullptr, (uint32_t)B, table, W, partials, L);
    }
    DMV_TRY(check_launch("table observable partials"));
    return launch_sum(partials, nblk, out);
}

template <class V>
__global__ void __launch_bounds__(RT) k_inner(const V *__restrict__ a, const V *__restrict__ b, uint32_t B,
                                              double *partials, StateLayout L)
{
    __shared__ double red[RT / 64];
    const uint64_t base = elem((uint64_t)blockIdx.x << B, L);
    const uint32_t span = 1u << B;
    double re = 0.0, im = 0.0;
    for (uint32_t i = threadIdx.x; i < span; i += RT) {
        const V x = a[base ^ i], y = b[base ^ i];
        re += (double)x.x * y.x + (double)x.y * y.y;
        im += (double)x.x * y.y - (double)x.y * y.x;
    }
    re = block_sum<RT>(re, red);
    im = block_sum<RT>(im, red);
    if (threadIdx.x == 0) {
        partials[2 * (uint64_t)blockIdx.x] = re;
        partials[2 * (uint64_t)blockIdx.x + 1] = im;
    }
}

int launch_inner(int dtype, const void *a, const void *b, int n, StateLayout L, double *partials, double *out2)
{
    const int B = n < kBlockBits ? n : kBlockBits;
    const uint64_t nblk = uint64_t(1) << (n - B);
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_inner<c64>, dim3((uint32_t)nblk), dim3(RT), 0, stream(), (const c64 *)a, (const c64 *)b,
                           (uint32_t)B, partials, L);
    else
        hipLaunchKernelGGL(k_inner<c128>, dim3((uint32_t)nblk), dim3(RT), 0, stream(), (const c128 *)a,
                           (const c128 *)b, (uint32_t)B, partials, L);
    DMV_TRY(check_launch("inner product partials"));
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(RT), 0, stream(), partials, nblk, 2u, 2u, out2);
    return check_launch("inner product final");
}

// One workgroup per row b: out[slot_map ? slot_map[b] : b] = sum_t
// partials[b * n_tiles + t], in a fixed order.
__global__ void __launch_bounds__(RT) k_slot_sums(const double *__restrict__ partials, uint64_t n_tiles,
                                                  const int *__restrict__ slot_map, double *out)
{
    __shared__ double red[RT / 64];
    const double *p = partials + (uint64_t)blockIdx.x * n_tiles;
    double acc = 0.0;
    for (uint64_t t = threadIdx.x; t < n_tiles; t += RT)
        acc += p[t];
    acc = block_sum<RT>(acc, red);
    if (threadIdx.x == 0)
        out[slot_map ? slot_map[blockIdx.x] : (int)blockIdx.x] = acc;
}

int launch_slot_sums(const double *partials, int n_rows, uint64_t n_tiles, const int *slot_map, double *out)
{
    if (n_rows <= 0)
        return DMV_OK;
    hipLaunchKernelGGL(k_slot_sums, dim3((uint32_t)n_rows), dim3(RT), 0, stream(), partials, n_tiles, slot_map, out);
    return check_launch("slot sums");
}

// ----------------------------------------------------------------- sampler --
// Inclusive prefix of the block masses, one workgroup, fixed order.
__global__ void __launch_bounds__(RT) k_block_scan(const double *__restrict__ in, uint64_t count, double *cum)
{
    __shared__ double tsum[RT];
    const uint64_t seg = (count + RT - 1) / RT;
    const uint64_t b0 = threadIdx.x * seg;
    const uint64_t b1 = b0 + seg < count ? b0 + seg : count;
    double s = 0.0;
    for (uint64_t i = b0; i < b1; i++)
        s += in[i];
    tsum[threadIdx.x] = s;
    __syncthreads();
    double off = 0.0;
    for (uint32_t t = 0; t < threadIdx.x; t++)
        off += tsum[t];
    for (uint64_t i = b0; i < b1; i++) {
        off += in[i];
        cum[i] = off;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One wave per sample: binary search over the block prefix, then a wave scan
// of the chosen block, 64 amplitudes at a time.
template <class V>
__global__ void __launch_bounds__(RT) k_sample(const V *__restrict__ s, uint32_t bb, uint64_t nblk,
                                               const double *__restrict__ cum, int64_t k, uint64_t seed,
                                               const double *__restrict__ u_in, int64_t *__restrict__ out,
                                               StateLayout L)
{
    const int64_t w = ((int64_t)blockIdx.x * RT + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= k)
        return;
    double u;
    if (u_in) {
        u = u_in[w];
    } else {
        const uint64_t z = mix64(seed + (uint64_t)(w + 1) * 0x9E3779B97F4A7C15ull);
        u = (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);
    }
    const double xsi = u * cum[nblk - 1];
    uint64_t lo = 0, hi = nblk - 1;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (xsi <= cum[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    const uint64_t b = lo;
    double base = b > 0 ? cum[b - 1] : 0.0;
    const uint32_t span = 1u << bb;
    const uint64_t first = b << bb;
    const uint64_t eb = elem(first, L);
    int64_t result = -1, last_nz = -1;
    for (uint32_t c0 = 0; c0 < span; c0 += 64) {
        const uint32_t j = c0 + lane;
        const double p = j < span ? norm2(s[eb ^ j]) : 0.0;
        double sc = p;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double t = __shfl_up(sc, off, 64);
            if (lane >= off)
                sc += t;
        }
        const double c = base + sc;
        const uint64_t hit = __ballot(p > 0.0 && xsi <= c);
        const uint64_t nz = __ballot(p > 0.0);
        if (nz)
            last_nz = (int64_t)(first + c0 + 63 - __clzll(nz));
        if (hit) {
            result = (int64_t)(first + c0 + __ffsll((long long)hit) - 1);
            break;
        }
        base = __shfl(c, 63, 64);
    }
    if (result < 0)
        result = last_nz;
    if (lane == 0)
        out[w] = result;
}

int launch_sample(int dtype, const void *state, int n, StateLayout L, int64_t k, uint64_t seed, const double *u,
                  double *block_cum, double *partials, int64_t *out)
{
    const int bb = n < kBlockBits ? n : kBlockBits;
    const uint64_t nblk = uint64_t(1) << (n - bb);
    DMV_TRY(launch_block_norms(dtype, state, n, L, bb, partials));
    hipLaunchKernelGGL(k_block_scan, dim3(1), dim3(RT), 0, stream(), partials, nblk, block_cum);
    DMV_TRY(check_launch("block scan"));
    const uint64_t waves_per_block = RT / 64;
    const uint64_t blocks = ((uint64_t)k + waves_per_block - 1) / waves_per_block;
    DMV_CHECK_ARG(blocks <= 0x7FFFFFFFull, "too many samples");
    if (dtype == DMV_C64)
        hipLaunchKernelGGL(k_sample<c64>, dim3((uint32_t)blocks), dim3(RT), 0, stream(), (const c64 *)state,
                           (uint32_t)bb, nblk, block_cum, k, seed, u, out, L);
    else
        hipLaunchKernelGGL(k_sample<c128>, dim3((uint32_t)blocks), dim3(RT), 0, stream(), (const c128 *)state,
                           (uint32_t)bb, nblk, block_cum, k, seed, u, out, L);
    return check_launch("sampler");
}

}  // namespace dmv
