// rng.hip -- bulk stream fills: raw words, U01 doubles and normals.
//
// Reference: CounterEngine batch generation (include/vsmc/rng/counter.hpp:
// 357-390), u01_distribution (u01_distribution.hpp:218-233) and the batch
// normal_distribution (normal_distribution.hpp:222-267).
#include "internal.hpp"
#include "../../include/vsmch/device_rng.hpp"

namespace vsmch {

enum { kPhilox = 0, kThreefry = 1 };

// Four consecutive words starting at stream word w (any alignment).
__device__ inline void philox_words4(uint32_t key, unsigned long long w,
    uint32_t out[4])
{
    const unsigned long long b = w / 4 + 1;
    const int r = static_cast<int>(w % 4);
    const u32x4 x = philox_block(key, b);
    if (r == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            out[i] = x.v[i];
        return;
    }
    const u32x4 y = philox_block(key, b + 1);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        out[i] = (r + i < 4) ? x.v[r + i] : y.v[r + i - 4];
}

__device__ inline void threefry_words4(unsigned long long key,
    unsigned long long w, unsigned long long out[4])
{
    const unsigned long long b = w / 4 + 1;
    const int r = static_cast<int>(w % 4);
    const u64x4 x = threefry_block(key, b);
    if (r == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
            out[i] = x.v[i];
        return;
    }
    const u64x4 y = threefry_block(key, b + 1);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        out[i] = (r + i < 4) ? x.v[r + i] : y.v[r + i - 4];
}

// mode 0: raw words; mode 1: U01 doubles.
template <int kKind, int kMode>
__global__ void __launch_bounds__(256) k_stream_fill(unsigned long long key,
    unsigned long long offset, long long n, void *out)
{
    const long long g = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long k = 4 * g;
    if (k >= n)
        return;
    const unsigned long long w = offset + static_cast<unsigned long long>(k);
    if (kKind == kPhilox) {
        uint32_t x[4];
        philox_words4(static_cast<uint32_t>(key), w, x);
        if (kMode == 0) {
            uint32_t *o = static_cast<uint32_t *>(out);
            if (k + 4 <= n) {
                *reinterpret_cast<uint4 *>(o + k) = make_uint4(x[0], x[1], x[2], x[3]);
            } else {
                for (int i = 0; i < 4 && k + i < n; ++i)
                    o[k + i] = x[i];
            }
        } else {
            double *o = static_cast<double *>(out);
            if (k + 4 <= n) {
                double2 a = make_double2(u01_u32d(x[0]), u01_u32d(x[1]));
                double2 b = make_double2(u01_u32d(x[2]), u01_u32d(x[3]));
                *reinterpret_cast<double2 *>(o + k) = a;
                *reinterpret_cast<double2 *>(o + k + 2) = b;
            } else {
                for (int i = 0; i < 4 && k + i < n; ++i)
                    o[k + i] = u01_u32d(x[i]);
            }
        }
    } else {
        unsigned long long x[4];
        threefry_words4(key, w, x);
        if (kMode == 0) {
            unsigned long long *o = static_cast<unsigned long long *>(out);
            for (int i = 0; i < 4 && k + i < n; ++i)
                o[k + i] = x[i];
        } else {
            double *o = static_cast<double *>(out);
            if (k + 4 <= n) {
                double2 a = make_double2(u01_u64d(x[0]), u01_u64d(x[1]));
                double2 b = make_double2(u01_u64d(x[2]), u01_u64d(x[3]));
                *reinterpret_cast<double2 *>(o + k) = a;
                *reinterpret_cast<double2 *>(o + k + 2) = b;
            } else {
                for (int i = 0; i < 4 && k + i < n; ++i)
                    o[k + i] = u01_u64d(x[i]);
            }
        }
    }
}

template <int kKind>
__device__ inline double u01_at(unsigned long long key, unsigned long long w)
{
    if (kKind == kPhilox)
        return u01_u32d(philox_word(static_cast<uint32_t>(key), w));
    return u01_u64d(threefry_word(key, w));
}

template <int kKind>
__device__ inline void u01_4(unsigned long long key, unsigned long long w,
    double out[4])
{
    if (kKind == kPhilox) {
        uint32_t x[4];
        philox_words4(static_cast<uint32_t>(key), w, x);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            out[i] = u01_u32d(x[i]);
    } else {
        unsigned long long x[4];
        threefry_words4(key, w, x);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            out[i] = u01_u64d(x[i]);
    }
}

// Batch normal_distribution: per chunk of 1024 (or the remainder k),
// nu = k / 2 radii then nu angles; r[i] = (sd * sqrt(-2 log u_i)) *
// sin(2 pi u_{nu+i}) + mean, r[nu+i] the same with cos. Odd n: the last
// element comes from two further words (normal_distribution.hpp:259-266).
// Each thread does four pairs.
template <int kKind>
__global__ void __launch_bounds__(256) k_normal_batch(unsigned long long key,
    unsigned long long offset, long long n, double mean, double sd,
    double *out)
{
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long chunk = q / 128;   // 512 pairs per full chunk, 4 per thread
    const long long c0 = chunk * 1024;
    if (c0 >= n)
        return;
    const long long kc = (n - c0) < 1024 ? (n - c0) : 1024;
    const long long nu = kc / 2;
    const long long i0 = (q % 128) * 4;
    if ((n % 2) != 0 && c0 + kc == n && i0 == 0) {
        // odd total: r[n-1] from words offset+n, offset+n+1
        const double u = u01_at<kKind>(key, offset + n);
        const double v = u01_at<kKind>(key, offset + n + 1);
        const double r = rn_mul(sd, sqrt(rn_mul(-2.0, log(u))));
        out[n - 1] = rn_add(mean,
            rn_mul(r, cos(rn_mul(two_pi<double>(), v))));
    }
    if (i0 >= nu)
        return;
    const unsigned long long wr = offset + c0 + i0;
    const unsigned long long wa = offset + c0 + nu + i0;
    double ur[4], ua[4];
    if (i0 + 4 <= nu) {
        u01_4<kKind>(key, wr, ur);
        u01_4<kKind>(key, wa, ua);
    } else {
        for (int t = 0; t < 4; ++t) {
            ur[t] = (i0 + t < nu) ? u01_at<kKind>(key, wr + t) : 0.5;
            ua[t] = (i0 + t < nu) ? u01_at<kKind>(key, wa + t) : 0.5;
        }
    }
    for (int t = 0; t < 4; ++t) {
        const long long i = i0 + t;
        if (i >= nu)
            break;
        double s = log(ur[t]);
        s = rn_mul(-2.0, s);
        s = sqrt(s);
        const double a = rn_mul(two_pi<double>(), ua[t]);
        double sn, cs;
        sincos(a, &sn, &cs);
        s = rn_mul(sd, s);
        out[c0 + i] = rn_add(rn_mul(s, sn), mean);
        out[c0 + nu + i] = rn_add(rn_mul(s, cs), mean);
    }
}

// Scalar NormalDistribution over consecutive words: pair p gives
// mean + sd * (r cos a) and mean + sd * (r sin a) (normal_distribution.hpp:
// 199-216).
template <int kKind>
__global__ void __launch_bounds__(256) k_normal_scalar(unsigned long long key,
    unsigned long long offset, long long n, double mean, double sd,
    double *out)
{
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (2 * p >= n)
        return;
    const double ua = u01_at<kKind>(key, offset + 2 * p);
    const double ub = u01_at<kKind>(key, offset + 2 * p + 1);
    const double r = sqrt(rn_mul(-2.0, log(ua)));
    const double a = rn_mul(two_pi<double>(), ub);
    double sn, cs;
    sincos(a, &sn, &cs);
    out[2 * p] = rn_add(mean, rn_mul(sd, rn_mul(r, cs)));
    if (2 * p + 1 < n)
        out[2 * p + 1] = rn_add(mean, rn_mul(sd, rn_mul(r, sn)));
}

int stream_fill(int kind, unsigned long long key, unsigned long long offset,
    long long n, void *out, int mode, hipStream_t st)
{
    if (n <= 0)
        return kOk;
    const long long threads = (n + 3) / 4;
    const dim3 grid(static_cast<unsigned int>((threads + 255) / 256));
    if (kind == kPhilox) {
        if (mode == 0)
            hipLaunchKernelGGL((k_stream_fill<kPhilox, 0>), grid, dim3(256), 0,
                st, key, offset, n, out);
        else
            hipLaunchKernelGGL((k_stream_fill<kPhilox, 1>), grid, dim3(256), 0,
                st, key, offset, n, out);
    } else if (kind == kThreefry) {
        if (mode == 0)
            hipLaunchKernelGGL((k_stream_fill<kThreefry, 0>), grid, dim3(256),
                0, st, key, offset, n, out);
        else
            hipLaunchKernelGGL((k_stream_fill<kThreefry, 1>), grid, dim3(256),
                0, st, key, offset, n, out);
    } else {
        set_error("unknown rng kind");
        return kErrInvalid;
    }
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int normal_fill(int kind, unsigned long long key, unsigned long long offset,
    long long n, double mean, double sd, int form, double *out,
    hipStream_t st)
{
    if (n <= 0)
        return kOk;
    if (kind != kPhilox && kind != kThreefry) {
        set_error("unknown rng kind");
        return kErrInvalid;
    }
    if (form == 0) {
        const long long chunks = (n + 1023) / 1024;
        const long long threads = chunks * 128;
        const dim3 grid(static_cast<unsigned int>((threads + 255) / 256));
        if (kind == kPhilox)
            hipLaunchKernelGGL(k_normal_batch<kPhilox>, grid, dim3(256), 0, st,
                key, offset, n, mean, sd, out);
        else
            hipLaunchKernelGGL(k_normal_batch<kThreefry>, grid, dim3(256), 0,
                st, key, offset, n, mean, sd, out);
    } else {
        const long long pairs = (n + 1) / 2;
        const dim3 grid(static_cast<unsigned int>((pairs + 255) / 256));
        if (kind == kPhilox)
            hipLaunchKernelGGL(k_normal_scalar<kPhilox>, grid, dim3(256), 0,
                st, key, offset, n, mean, sd, out);
        else
            hipLaunchKernelGGL(k_normal_scalar<kThreefry>, grid, dim3(256), 0,
                st, key, offset, n, mean, sd, out);
    }
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

} // namespace vsmch
