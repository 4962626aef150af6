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
