// device_rng.hpp -- counter-based streams of vSMC on gfx950, stateless.
//
// A vSMC CounterEngine (include/vsmc/rng/counter.hpp:288-496) seeded with
// s has key {s, 0, ...} and counter 0, increments the counter before it
// encrypts a block (philox.hpp:263-269, threefry.hpp:329-336), and hands
// out the block's words in order. Word k of such a stream is therefore word
// k % 4 of the block whose counter is k / 4 + 1. The functions here compute
// any word directly from (key, k), so a particle's stream needs no state in
// HBM: the position is (step, draw) arithmetic.
//
// Everything is __host__ __device__ so the same code serves kernels and the
// host-side parts of the library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsmch {

// ------------------------------------------------------------ Philox4x32-10
// include/vsmc/rng/philox.hpp:76-93 (constants), :208-228 (round),
// :263-295 (10 rounds, key bumped before rounds 2..10).
struct u32x4 {
    uint32_t v[4];
};

__host__ __device__ inline u32x4 philox4x32_10(
    uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
    uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u;
    const uint32_t M1 = 0xCD9E8D57u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // each product whole (one v_mad_u64_u32 on the device instead of a
        // v_mul_hi_u32 and a v_mul_lo_u32)
        const uint64_t p1 = static_cast<uint64_t>(c2) * M1;
        const uint64_t p0 = static_cast<uint64_t>(c0) * M0;
        c0 = static_cast<uint32_t>(p1 >> 32) ^ k0 ^ c1;
        c1 = static_cast<uint32_t>(p1);
        c2 = static_cast<uint32_t>(p0 >> 32) ^ k1 ^ c3;
        c3 = static_cast<uint32_t>(p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    u32x4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// Block b (1-based counter value) of the Philox stream with key {key0, 0}.
__host__ __device__ inline u32x4 philox_block(uint32_t key0, uint64_t b)
{
    return philox4x32_10(static_cast<uint32_t>(b),
        static_cast<uint32_t>(b >> 32), 0u, 0u, key0, 0u);
}

// Word k (0-based) of the stream seeded with key0.
__host__ __device__ inline uint32_t philox_word(uint32_t key0, uint64_t k)
{
    const u32x4 blk = philox_block(key0, k / 4 + 1);
    // a select, not blk.v[k % 4]: a run-time index puts blk in scratch
    const unsigned r = static_cast<unsigned>(k & 3);
    return r == 0 ? blk.v[0] : r == 1 ? blk.v[1] : r == 2 ? blk.v[2] : blk.v[3];
}

// ---------------------------------------------------------- Threefry4x64-20
// include/vsmc/rng/threefry.hpp:71-82 (parity constant), :120-135
// (rotations), :148-298 (rounds and key injection), :304-367 (generator).
struct u64x4 {
    uint64_t v[4];
};

__host__ __device__ inline uint64_t rotl64c(uint64_t x, int r)
{
    return (x << r) | (x >> (64 - r));
}

__host__ __device__ inline u64x4 threefry4x64_20(
    const uint64_t ctr[4], const uint64_t key[4])
{
    uint64_t ks[5];
    ks[4] = 0x1BD11BDAA9FC1A22ull;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ks[i] = key[i];
        ks[4] ^= key[i];
    }
    uint64_t x0 = ctr[0] + ks[0];
    uint64_t x1 = ctr[1] + ks[1];
    uint64_t x2 = ctr[2] + ks[2];
    uint64_t x3 = ctr[3] + ks[3];
    // Rounds come in groups of four: (0,1)(2,3) then (0,3)(2,1) alternate.
    const int R0[8] = {14, 52, 23, 5, 25, 46, 58, 32};
    const int R1[8] = {16, 57, 40, 37, 33, 12, 22, 32};
#pragma unroll
    for (int g = 0; g < 5; ++g) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = (4 * g + q) % 8;
            if (q % 2 == 0) {
                x0 += x1;
                x1 = rotl64c(x1, R0[r]);
                x1 ^= x0;
                x2 += x3;
                x3 = rotl64c(x3, R1[r]);
                x3 ^= x2;
            } else {
                x0 += x3;
                x3 = rotl64c(x3, R0[r]);
                x3 ^= x0;
                x2 += x1;
                x1 = rotl64c(x1, R1[r]);
                x1 ^= x2;
            }
        }
        const int inc = g + 1;
        x0 += ks[(inc + 0) % 5];
        x1 += ks[(inc + 1) % 5];
        x2 += ks[(inc + 2) % 5];
        x3 += ks[(inc + 3) % 5] + static_cast<uint64_t>(inc);
    }
    u64x4 out;
    out.v[0] = x0;
    out.v[1] = x1;
    out.v[2] = x2;
    out.v[3] = x3;
    return out;
}

__host__ __device__ inline u64x4 threefry_block(uint64_t key0, uint64_t b)
{
    const uint64_t ctr[4] = {b, 0, 0, 0};
    const uint64_t key[4] = {key0, 0, 0, 0};
    return threefry4x64_20(ctr, key);
}

__host__ __device__ inline uint64_t threefry_word(uint64_t key0, uint64_t k)
{
    const u64x4 blk = threefry_block(key0, k / 4 + 1);
    return blk.v[k % 4];
}

// ---------------------------------------------------------------------- U01
// u01_co of include/vsmc/rng/u01.hpp:126-147: [0, 1) with the full width of
// the word or the mantissa, whichever is smaller.
__host__ __device__ inline double u01_u32d(uint32_t u)
{
    return static_cast<double>(u) * (1.0 / 4294967296.0);
}

__host__ __device__ inline double u01_u64d(uint64_t u)
{
    return static_cast<double>(u >> 11) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline float u01_u32f(uint32_t u)
{
    return static_cast<float>(u >> 8) * (1.0f / 16777216.0f);
}

// 2 pi as const_pi_2<T>() (include/vsmc/math/constants.hpp:66).
template <typename T>
__host__ __device__ inline T two_pi();
template <>
__host__ __device__ inline double two_pi<double>()
{
    return 6.283185307179586476925286766559005768394;
}
template <>
__host__ __device__ inline float two_pi<float>()
{
    return 6.283185307179586476925286766559005768394f;
}

template <typename T>
__host__ __device__ inline T u01_from_u32(uint32_t u);
template <>
__host__ __device__ inline double u01_from_u32<double>(uint32_t u)
{
    return u01_u32d(u);
}
template <>
__host__ __device__ inline float u01_from_u32<float>(uint32_t u)
{
    return u01_u32f(u);
}

// Box-Muller pair as NormalDistribution::generate draws it
// (include/vsmc/rng/normal_distribution.hpp:199-216): the first word is the
// radius, the second the angle; the first call returns r cos, the cached
// second call returns r sin.
template <typename T>
struct normal_pair {
    T z0;
    T z1;
};

// Round-to-nearest add and multiply that the compiler may not contract into
// an FMA: the reference's host build evaluates a * b + c with two roundings.
// HIP's __dadd_rn / __dmul_rn are plain + and *, so the operations here
// carry no contraction permission of their own (HIP's default contraction
// mode honours the pragma; translation units of the library are also built
// with -ffp-contract=off).
__device__ inline double rn_add(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ inline float rn_add(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}
__device__ inline double rn_mul(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}
__device__ inline float rn_mul(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}

__device__ inline void sincos_t(double a, double *s, double *c)
{
    sincos(a, s, c);
}
__device__ inline void sincos_t(float a, float *s, float *c)
{
    sincosf(a, s, c);
}

template <typename T>
__device__ inline normal_pair<T> box_muller(uint32_t w_radius, uint32_t w_angle)
{
    const T ua = u01_from_u32<T>(w_radius);
    const T ub = u01_from_u32<T>(w_angle);
    const T r = sqrt(static_cast<T>(-2) * log(ua));
    const T a = two_pi<T>() * ub;
    T s, c;
    sincos_t(a, &s, &c);
    normal_pair<T> p;
    p.z0 = r * c;
    p.z1 = r * s;
    return p;
}

// ---------------------------------------------------------------- streams
// Seed::get() hands out s+1, s+2, ..., UINT_MAX, 1, 2, ... for the default
// SeedGenerator<..., unsigned> (include/vsmc/rng/seed.hpp:118-123,
// :172-178); particle i of a Particle<T> built after Seed::set(s) has key
// seed_key(s, i) and the resampling engine seed_key(s, N)
// (rng_set.hpp:88-115, particle.hpp:60-67).
__host__ __device__ inline uint32_t seed_key(uint32_t s, uint64_t k)
{
    const uint64_t m = 0xFFFFFFFFull;
    return static_cast<uint32_t>((static_cast<uint64_t>(s % m) + k) % m) + 1u;
}

// A particle's engine: key plus the next unread word. Device functors use it
// as sp.rng(): each call of next() returns the next word of the particle's
// stream exactly as CounterEngine::operator() would.
struct ParticleStream {
    uint32_t key;
    uint64_t word;
    uint32_t blk[4];
    uint64_t blk_id;

    __host__ __device__ ParticleStream(uint32_t k, uint64_t w)
        : key(k), word(w), blk_id(0)
    {
    }

    __host__ __device__ uint32_t next()
    {
        const uint64_t b = word / 4 + 1;
        if (b != blk_id) {
            const u32x4 r = philox_block(key, b);
            blk[0] = r.v[0];
            blk[1] = r.v[1];
            blk[2] = r.v[2];
            blk[3] = r.v[3];
            blk_id = b;
        }
        const uint32_t out = blk[word % 4];
        ++word;
        return out;
    }
};

} // namespace vsmch
