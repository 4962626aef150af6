// vsmch/random_walk.hpp -- random-walk Metropolis-Hastings for device
// functors: RandomWalk, RandomWalkG, NormalProposal and NormalMVProposal
// (include/vsmc/rng/random_walk.hpp:49-356, :358-555) with the reference's
// draw order, so an eval_sp that moves a particle with them consumes the
// same stream words as the reference functor (up to libm ulps in exp / log
// and the Box-Muller transcendentals).
//
// Differences from the reference: the dimension is a compile-time constant
// (the state lives in registers), and NormalMVProposal multiplies by the
// Cholesky factor in row order, left to right (the reference calls an
// external BLAS tpmv, whose summation order is its own).
#pragma once

#include "distribution.hpp"

#include <cmath>
#include <cstddef>

namespace vsmch {

namespace rw {

// y = x + z, log q-ratio 0 (random_walk.hpp:378-384)
template <typename T>
__device__ inline T q_free(T x, T &y, T z)
{
    y = rn_add(x, z);
    return 0;
}

// support (a, inf): y = a + (x - a) exp(z), log q-ratio z (:386-392)
template <typename T>
__device__ inline T q_lower(T x, T &y, T z, T a)
{
    y = rn_add(a, rn_mul(x - a, static_cast<T>(exp(z))));
    return z;
}

// support (-inf, b): y = b - (b - x) exp(z), log q-ratio z (:394-400)
template <typename T>
__device__ inline T q_upper(T x, T &y, T z, T b)
{
    y = b - rn_mul(b - x, static_cast<T>(exp(z)));
    return z;
}

// support (a, b): logit random walk (:402-410)
template <typename T>
__device__ inline T q_both(T x, T &y, T z, T a, T b)
{
    const T r = rn_mul(static_cast<T>(exp(z)), x - a) / (b - x);
    y = rn_add(a, rn_mul(b, r)) / rn_add(static_cast<T>(1), r);
    return static_cast<T>(log((y - a) / (x - a) * (b - y) / (b - x)));
}

// flag: 2 (lower bound) + 1 (upper bound), as the reference's constructor
template <typename T>
__device__ inline T q_apply(unsigned flag, T x, T &y, T z, T a, T b)
{
    switch (flag) {
    case 0: return q_free(x, y, z);
    case 1: return q_upper(x, y, z, b);
    case 2: return q_lower(x, y, z, a);
    case 3: return q_both(x, y, z, a, b);
    default: return 0;
    }
}

template <typename T>
__host__ __device__ inline unsigned bound_flag(T a, T b)
{
    const unsigned lower = (a > -INFINITY && a < INFINITY) ? 1u : 0u;
    const unsigned upper = (b > -INFINITY && b < INFINITY) ? 1u : 0u;
    return (lower << 1) + upper;
}

} // namespace rw

// NormalProposal<RealType> (random_walk.hpp:414-452): z ~ N(0, stddev) from
// a NormalDistribution that keeps its cached Box-Muller half between calls.
template <typename RealType>
class NormalProposal {
  public:
    using result_type = RealType;

    __host__ __device__ explicit NormalProposal(RealType stddev = 1,
        RealType a = -INFINITY, RealType b = INFINITY)
        : rnorm_(0, stddev), a_(a), b_(b), flag_(rw::bound_flag(a, b))
    {
    }

    __host__ __device__ RealType a() const { return a_; }
    __host__ __device__ RealType b() const { return b_; }

    // proposes y and returns log(q(y, x) / q(x, y))
    template <typename RNG>
    __device__ RealType operator()(RNG &rng, std::size_t, const RealType *x,
        RealType *y)
    {
        const RealType z = rnorm_(rng);
        return rw::q_apply(flag_, *x, *y, z, a_, b_);
    }

  private:
    NormalDistribution<RealType> rnorm_;
    RealType a_;
    RealType b_;
    unsigned flag_;
};

// NormalMVProposal<RealType, Dim> (random_walk.hpp:456-555): z = L n with n
// Dim standard normals (the scalar form, fewer than 100 components,
// normal_distribution.hpp:126-136) and L the lower-triangular Cholesky factor
// packed row by row (NULL: the identity); each component then takes its own
// bounded transform and the log q-ratios add.
template <typename RealType, std::size_t Dim>
class NormalMVProposal {
  public:
    using result_type = RealType;

    __host__ __device__ explicit NormalMVProposal(const RealType *chol = nullptr,
        const RealType *a = nullptr, const RealType *b = nullptr)
        : rnorm_(0, 1), has_chol_(chol != nullptr)
    {
        for (std::size_t k = 0; k != Dim * (Dim + 1) / 2; ++k)
            chol_[k] = chol ? chol[k] : 0;
        for (std::size_t i = 0; i != Dim; ++i) {
            a_[i] = a ? a[i] : -INFINITY;
            b_[i] = b ? b[i] : INFINITY;
            flag_[i] = rw::bound_flag(a_[i], b_[i]);
        }
    }

    __host__ __device__ static constexpr std::size_t dim() { return Dim; }

    template <typename RNG>
    __device__ RealType operator()(RNG &rng, std::size_t, const RealType *x,
        RealType *y)
    {
        RealType z[Dim];
        for (std::size_t i = 0; i != Dim; ++i)
            z[i] = rnorm_(rng);
        if (has_chol_) {
            // r = L r in place, last row first (tpmv, RowMajor Lower NoTrans)
            for (std::size_t i = Dim; i-- > 0;) {
                RealType s = 0;
                for (std::size_t j = 0; j <= i; ++j)
                    s = rn_add(s, rn_mul(chol_[i * (i + 1) / 2 + j], z[j]));
                z[i] = s;
            }
        }
        RealType q = 0;
        for (std::size_t i = 0; i != Dim; ++i)
            q = rn_add(q, rw::q_apply(flag_[i], x[i], y[i], z[i], a_[i], b_[i]));
        return q;
    }

  private:
    NormalDistribution<RealType> rnorm_;
    RealType chol_[Dim * (Dim + 1) / 2];
    RealType a_[Dim];
    RealType b_[Dim];
    unsigned flag_[Dim];
    bool has_chol_;
};

// RandomWalk<RealType, Dim> (random_walk.hpp:49-191): accept the proposal y
// when log U < log gamma(y) - log gamma(x) + log q-ratio; ltx (if not null)
// holds log gamma(x) between calls. log_target(dim, x) and
// proposal(rng, dim, x, y) are device callables.
template <typename RealType, std::size_t Dim>
class RandomWalk {
  public:
    using result_type = RealType;

    __host__ __device__ static constexpr std::size_t dim() { return Dim; }

    // one-step update (:97-117); returns the acceptance count
    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(RNG &rng, RealType *x, RealType *ltx,
        LogTarget &&log_target, Proposal &&proposal)
    {
        U01Distribution<RealType> u01;
        RealType y[Dim];
        const RealType q = proposal(rng, Dim, x, y);
        const RealType s = ltx == nullptr ? log_target(Dim, x) : *ltx;
        const RealType t = log_target(Dim, y);
        const RealType p = rn_add(t - s, q);
        const RealType u = static_cast<RealType>(log(u01(rng)));
        if (u < p) {
            for (std::size_t i = 0; i != Dim; ++i)
                x[i] = y[i];
            if (ltx != nullptr)
                *ltx = t;
            return 1;
        }
        return 0;
    }

    // one-step update of x[idx .. idx + Dim) of a state of length m
    // (:126-148); log_target sees all m components
    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(RNG &rng, std::size_t m, std::size_t idx,
        RealType *x, RealType *ltx, LogTarget &&log_target, Proposal &&proposal)
    {
        U01Distribution<RealType> u01;
        RealType xo[Dim], y[Dim];
        for (std::size_t i = 0; i != Dim; ++i)
            xo[i] = x[idx + i];
        const RealType q = proposal(rng, Dim, xo, y);
        const RealType s = ltx == nullptr ? log_target(m, x) : *ltx;
        for (std::size_t i = 0; i != Dim; ++i)
            x[idx + i] = y[i];
        const RealType t = log_target(m, x);
        const RealType p = rn_add(t - s, q);
        const RealType u = static_cast<RealType>(log(u01(rng)));
        if (u < p) {
            if (ltx != nullptr)
                *ltx = t;
            return 1;
        }
        for (std::size_t i = 0; i != Dim; ++i)
            x[idx + i] = xo[i];
        return 0;
    }

    // n-step update (:151-167)
    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(std::size_t n, RNG &rng, RealType *x,
        RealType *ltx, LogTarget &&log_target, Proposal &&proposal)
    {
        std::size_t acc = 0;
        RealType s = ltx == nullptr ? log_target(Dim, x) : *ltx;
        for (std::size_t i = 0; i != n; ++i)
            acc += (*this)(rng, x, &s, log_target, proposal);
        if (ltx != nullptr)
            *ltx = s;
        return acc;
    }

    // n-step update of a block (:170-188)
    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(std::size_t n, RNG &rng, std::size_t m,
        std::size_t idx, RealType *x, RealType *ltx, LogTarget &&log_target,
        Proposal &&proposal)
    {
        std::size_t acc = 0;
        RealType s = ltx == nullptr ? log_target(m, x) : *ltx;
        for (std::size_t i = 0; i != n; ++i)
            acc += (*this)(rng, m, idx, x, &s, log_target, proposal);
        if (ltx != nullptr)
            *ltx = s;
        return acc;
    }
};

// RandomWalkG<RealType, DimX, DimG> (random_walk.hpp:192-356): RandomWalk
// that also keeps a test function g(x) of DimG values, written by
// log_target(dim_x, dim_g, x, g) (g may be null) and copied out on accept.
template <typename RealType, std::size_t DimX, std::size_t DimG>
class RandomWalkG {
  public:
    using result_type = RealType;

    __host__ __device__ static constexpr std::size_t dim_x() { return DimX; }
    __host__ __device__ static constexpr std::size_t dim_g() { return DimG; }

    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(RNG &rng, RealType *x, RealType *ltx,
        RealType *g, LogTarget &&log_target, Proposal &&proposal)
    {
        U01Distribution<RealType> u01;
        RealType y[DimX], gy[DimG];
        const RealType q = proposal(rng, DimX, x, y);
        const RealType s = ltx == nullptr
            ? log_target(DimX, DimG, x, static_cast<RealType *>(nullptr))
            : *ltx;
        const RealType t =
            log_target(DimX, DimG, y, g == nullptr ? nullptr : gy);
        const RealType p = rn_add(t - s, q);
        const RealType u = static_cast<RealType>(log(u01(rng)));
        if (u < p) {
            for (std::size_t i = 0; i != DimX; ++i)
                x[i] = y[i];
            if (ltx != nullptr)
                *ltx = t;
            if (g != nullptr)
                for (std::size_t i = 0; i != DimG; ++i)
                    g[i] = gy[i];
            return 1;
        }
        return 0;
    }

    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(RNG &rng, std::size_t m, std::size_t idx,
        RealType *x, RealType *ltx, RealType *g, LogTarget &&log_target,
        Proposal &&proposal)
    {
        U01Distribution<RealType> u01;
        RealType xo[DimX], y[DimX], gy[DimG];
        for (std::size_t i = 0; i != DimX; ++i)
            xo[i] = x[idx + i];
        const RealType q = proposal(rng, DimX, xo, y);
        const RealType s = ltx == nullptr
            ? log_target(m, DimG, x, static_cast<RealType *>(nullptr))
            : *ltx;
        for (std::size_t i = 0; i != DimX; ++i)
            x[idx + i] = y[i];
        const RealType t =
            log_target(m, DimG, x, g == nullptr ? nullptr : gy);
        const RealType p = rn_add(t - s, q);
        const RealType u = static_cast<RealType>(log(u01(rng)));
        if (u < p) {
            if (ltx != nullptr)
                *ltx = t;
            if (g != nullptr)
                for (std::size_t i = 0; i != DimG; ++i)
                    g[i] = gy[i];
            return 1;
        }
        for (std::size_t i = 0; i != DimX; ++i)
            x[idx + i] = xo[i];
        return 0;
    }

    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(std::size_t n, RNG &rng, RealType *x,
        RealType *ltx, RealType *g, LogTarget &&log_target, Proposal &&proposal)
    {
        std::size_t acc = 0;
        RealType s = ltx == nullptr
            ? log_target(DimX, DimG, x, static_cast<RealType *>(nullptr))
            : *ltx;
        for (std::size_t i = 0; i != n; ++i)
            acc += (*this)(rng, x, &s, g, log_target, proposal);
        if (ltx != nullptr)
            *ltx = s;
        return acc;
    }

    template <typename RNG, typename LogTarget, typename Proposal>
    __device__ std::size_t operator()(std::size_t n, RNG &rng, std::size_t m,
        std::size_t idx, RealType *x, RealType *ltx, RealType *g,
        LogTarget &&log_target, Proposal &&proposal)
    {
        std::size_t acc = 0;
        RealType s = ltx == nullptr
            ? log_target(m, DimG, x, static_cast<RealType *>(nullptr))
            : *ltx;
        for (std::size_t i = 0; i != n; ++i)
            acc += (*this)(rng, m, idx, x, &s, g, log_target, proposal);
        if (ltx != nullptr)
            *ltx = s;
        return acc;
    }
};

} // namespace vsmch
