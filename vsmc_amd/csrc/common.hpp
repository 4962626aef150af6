// common.hpp -- shared device primitives of libvsmch (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace vsmch {

constexpr int kWave = 64;
constexpr int kBlock = 256;          // threads per tile block (4 waves)
constexpr int kItems = 16;           // elements per thread
constexpr int kTile = kBlock * kItems;  // 4096 elements per tile
constexpr int kWaves = kBlock / kWave;

// ----------------------------------------------------- wave / block scans
// __shfl_up for trivially copyable types of 4-byte multiples.
template <typename T>
__device__ inline T shfl_up_any(T v, int d)
{
    static_assert(sizeof(T) % 4 == 0, "4-byte multiple");
    constexpr int n = sizeof(T) / 4;
    union {
        T t;
        int w[n];
    } a, b;
    a.t = v;
#pragma unroll
    for (int i = 0; i < n; ++i)
        b.w[i] = __shfl_up(a.w[i], d, kWave);
    return b.t;
}

template <typename T>
__device__ inline T shfl_any(T v, int src)
{
    static_assert(sizeof(T) % 4 == 0, "4-byte multiple");
    constexpr int n = sizeof(T) / 4;
    union {
        T t;
        int w[n];
    } a, b;
    a.t = v;
#pragma unroll
    for (int i = 0; i < n; ++i)
        b.w[i] = __shfl(a.w[i], src, kWave);
    return b.t;
}

template <typename T>
__device__ inline T shfl_down_any(T v, int d)
{
    static_assert(sizeof(T) % 4 == 0, "4-byte multiple");
    constexpr int n = sizeof(T) / 4;
    union {
        T t;
        int w[n];
    } a, b;
    a.t = v;
#pragma unroll
    for (int i = 0; i < n; ++i)
        b.w[i] = __shfl_down(a.w[i], d, kWave);
    return b.t;
}

// Inclusive wave scan with an associative operator. Lanes read lane - d
// through ds_bpermute (__shfl_up); lanes below d keep their value.
template <typename T, typename Op>
__device__ inline T wave_inclusive_scan(T v, Op op)
{
    const int lane = threadIdx.x % kWave;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T u = shfl_up_any(v, d);
        if (lane >= d)
            v = op(u, v);
    }
    return v;
}

// Block-wide exclusive scan over kThreads threads. Returns the exclusive
// prefix of this thread; *total receives the block aggregate. `lds` needs
// kThreads / kWave entries. Contains __syncthreads (all threads must call).
template <int kThreads, typename T, typename Op>
__device__ inline T block_exclusive_scan_n(T v, Op op, T identity, T *lds,
    T *total)
{
    constexpr int nw = kThreads / kWave;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    T inc = wave_inclusive_scan(v, op);
    if (lane == kWave - 1)
        lds[wave] = inc;
    __syncthreads();
    T wpre = identity;
    T agg = identity;
#pragma unroll
    for (int w = 0; w < nw; ++w) {
        if (w < wave)
            wpre = op(wpre, lds[w]);
        agg = op(agg, lds[w]);
    }
    T ex = shfl_up_any(inc, 1);
    if (lane == 0)
        ex = identity;
    __syncthreads();
    *total = agg;
    return op(wpre, ex);
}

// The same over the kBlock threads of a tile block.
template <typename T, typename Op>
__device__ inline T block_exclusive_scan(T v, Op op, T identity, T *lds,
    T *total)
{
    return block_exclusive_scan_n<kBlock>(v, op, identity, lds, total);
}

// Exclusive scan of n values get(i) by one block of kThreads threads,
// put(i, prefix) for each; returns the total. Each thread reads kPer
// consecutive values per chunk of kThreads * kPer (all its loads in flight
// at once), so a chunk costs one round of loads and one block scan. All
// threads must call it; lds needs kThreads / kWave entries.
template <int kThreads, int kPer, typename T, typename Op, typename Get,
    typename Put>
__device__ __forceinline__ T block_scan_chunked_n(long long n, Get get, Put put,
    Op op, T identity, T *lds)
{
    T carry = identity;
    for (long long c0 = 0; c0 < n; c0 += static_cast<long long>(kThreads) * kPer) {
        const long long b0 = c0 + static_cast<long long>(threadIdx.x) * kPer;
        T v[kPer];
        T loc = identity;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = (b0 + k < n) ? get(b0 + k) : identity;
            loc = op(loc, v[k]);
        }
        T tot;
        T run = op(carry, block_exclusive_scan_n<kThreads>(loc, op, identity, lds, &tot));
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (b0 + k < n)
                put(b0 + k, run);
            run = op(run, v[k]);
        }
        carry = op(carry, tot);
    }
    return carry;
}

// The same over the kBlock threads of a tile block, kItems per thread.
template <typename T, typename Op, typename Get, typename Put>
__device__ __forceinline__ T block_scan_chunked(long long n, Get get, Put put,
    Op op, T identity, T *lds)
{
    return block_scan_chunked_n<kBlock, kItems>(n, get, put, op, identity, lds);
}

template <typename T, typename Op>
__device__ inline T wave_reduce(T v, Op op)
{
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1)
        v = op(v, shfl_down_any(v, d));
    return shfl_any(v, 0);
}

// Wave sums of V per-lane values (V a power of two, at most kWave) with
// V shuffles instead of V log2(kWave): each xor step over lane bit m halves
// the values a lane carries (the lower half of the lanes keeps the first
// half of its values, the upper half the second, each adding its partner's
// copy), then the remaining lane bits are summed plainly. Afterwards lane l
// holds the wave's total of value (l / (kWave / V)). The summation order is
// fixed by the lane layout, so results are deterministic.
template <int H, int V>
__device__ inline void wave_transpose_step(double (&v)[V], int lane)
{
    if constexpr (H >= 1) {
        constexpr int m = kWave / V * H;   // lane bit of this step
        const bool upper = (lane & m) != 0;
#pragma unroll
        for (int j = 0; j < H; ++j) {
            const double send = upper ? v[j] : v[j + H];
            const double keep = upper ? v[j + H] : v[j];
            v[j] = keep + __shfl_xor(send, m, kWave);
        }
        wave_transpose_step<H / 2>(v, lane);
    }
}

template <int V>
__device__ inline double wave_transpose_sum(double (&v)[V])
{
    static_assert(V >= 1 && V <= kWave && (V & (V - 1)) == 0, "V");
    wave_transpose_step<V / 2>(v, static_cast<int>(threadIdx.x % kWave));
    double r = v[0];
#pragma unroll
    for (int m = kWave / V / 2; m >= 1; m >>= 1)
        r += __shfl_xor(r, m, kWave);
    return r;
}

// Pins a loaded value in registers at this point, so that the compiler
// cannot move its load into the branch that uses it: a load under a branch
// is waited for before the branch ends, which serialises a thread's loads.
// Issue a batch of loads, then pin each one. (An empty asm statement: no
// instruction is emitted.)
template <typename T>
__device__ __forceinline__ void keep_loaded(T &v)
{
    asm volatile("" : "+v"(v));
}

// 2^(j/64), j = 0..63, as pairs (hi, lo): hi is 2^(j/64) rounded to
// double, lo the rounded remainder (computed to 80 digits; a CPU test
// recomputes both).
static __constant__ double kWexpTab[128] = {
    0x1.0000000000000p+0, 0x0.0p+0,
    0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56,
    0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55,
    0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57,
    0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54,
    0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59,
    0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54,
    0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54,
    0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55,
    0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55,
    0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54,
    0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55,
    0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54,
    0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55,
    0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55,
    0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54,
    0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55,
    0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54,
    0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54,
    0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56,
    0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55,
    0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58,
    0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59,
    0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56,
    0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56,
    0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54,
    0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55,
    0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54,
    0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54,
    0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54,
    0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54,
    0x1.6623882552225p+0, -0x1.bb60987591c34p-54,
    0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54,
    0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57,
    0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55,
    0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54,
    0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55,
    0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56,
    0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54,
    0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54,
    0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54,
    0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55,
    0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57,
    0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54,
    0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56,
    0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54,
    0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54,
    0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54,
    0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54,
    0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57,
    0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56,
    0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55,
    0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55,
    0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54,
    0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56,
    0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54,
    0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55,
    0x1.da9e603db3285p+0, 0x1.c2300696db532p-54,
    0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54,
    0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55,
    0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54,
    0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54,
    0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54,
    0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55
};

// e^x for the normalised weights W = e^(l - lmax) / sum (x <= 0), within
// 0.75 ulp of the real value (40 million samples over [-745, 0] against a
// long-double exp): x = (64 m + j) ln2 / 64 + r with |r| <= ln2 / 128 (two
// FMAs with ln2 / 64 split in two), e^r - 1 by its Taylor polynomial of
// degree 6 (truncation below 2^-64), then 2^(j/64) (1 + (e^r - 1)) from
// the (hi, lo) table and the exponent m added by a multiply while the
// result is normal. tab is kWexpTab or a copy of it (in LDS). Every kernel
// that forms the weights the resampler counts with, and the read-back of
// those weights, uses this one function, so the counts are exact for the
// numbers a caller reads. Below -746 the result is 0; NaN stays NaN.
__device__ __forceinline__ double wexp(double x, const double *tab)
{
    const double kd = rint(x * 0x1.71547652b82fep+6);
    double r = fma(-kd, 0x1.62e42fefa39efp-7, x);
    r = fma(-kd, 0x1.abc9e3b39803fp-62, r);
    double p = fma(r, 1.0 / 720.0, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    const double em1 = p * r;
    const int k = static_cast<int>(kd);
    const int j = 2 * (k & 63);
    const double th = tab[j], tl = tab[j + 1];
    const double y = th + fma(th, em1, tl);
    const int m = k >> 6;
    if (__builtin_expect(m < -1021, 0))
        return x > -746.0 ? ldexp(y, m) : (x != x ? x : 0.0);
    return y * __hiloint2double((m + 1023) << 20, 0);
}

__device__ __forceinline__ double wexp(double x) { return wexp(x, kWexpTab); }

// Copies kWexpTab into the LDS table tab (128 doubles); the block's threads
// must synchronise before reading it.
__device__ __forceinline__ void wexp_table_to_lds(double *tab)
{
    for (int i = threadIdx.x; i < 128; i += blockDim.x)
        tab[i] = kWexpTab[i];
}

// Block-wide reduction, result broadcast to all threads.
template <typename T, typename Op>
__device__ inline T block_reduce(T v, Op op, T *lds)
{
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    v = wave_reduce(v, op);
    if (lane == 0)
        lds[wave] = v;
    __syncthreads();
    T r = lds[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w)
        r = op(r, lds[w]);
    __syncthreads();
    return r;
}

struct OpAddD {
    __device__ double operator()(double a, double b) const { return a + b; }
};
struct OpAddL {
    __device__ long long operator()(long long a, long long b) const
    {
        return a + b;
    }
};
struct OpMinL {
    __device__ long long operator()(long long a, long long b) const
    {
        return a < b ? a : b;
    }
};
struct OpMinI {
    __device__ int operator()(int a, int b) const { return a < b ? a : b; }
};
struct OpMaxI {
    __device__ int operator()(int a, int b) const { return a > b ? a : b; }
};
struct OpMaxD {
    __device__ double operator()(double a, double b) const
    {
        return a > b ? a : b;
    }
};

// Load a tile of kTile doubles from `src` (n valid, tile-relative base) into
// blocked order (thread t owns elements t*kItems .. t*kItems+kItems-1),
// staging through LDS so the global reads are coalesced.
// LDS index with one pad slot per kItems entries, so that the blocked reads
// (stride kItems) do not all hit the same banks.
__device__ inline int lds_pad(int e) { return e + (e / kItems); }
constexpr int kTileLds = kTile + kTile / kItems;

// Stage the tile into LDS only (element k at lds_pad(k)), `pad` past n
// (base < n). The loads are unconditional (index clamped to n - 1) and all
// issued before the first LDS store, so they are in flight together: a
// branch or an interleaved store per element would serialise them.
template <typename Src>
__device__ __forceinline__ void load_tile_lds(Src src, long long base, long long n,
    double *lds_tile, double pad)
{
    const auto s = src.loaded();
    double v[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const long long g = base + j * kBlock + threadIdx.x;
        const double x = s(g < n ? g : n - 1);
        v[j] = g < n ? x : pad;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        lds_tile[lds_pad(j * kBlock + threadIdx.x)] = v[j];
    __syncthreads();
}

template <typename Src>
__device__ __forceinline__ void load_tile_blocked(Src src, long long base,
    long long n, double (&w)[kItems], double *lds_tile, double pad)
{
    load_tile_lds(src, base, n, lds_tile, pad);
#pragma unroll
    for (int j = 0; j < kItems; ++j)
        w[j] = lds_tile[lds_pad(threadIdx.x * kItems + j)];
    __syncthreads();
}

} // namespace vsmch
