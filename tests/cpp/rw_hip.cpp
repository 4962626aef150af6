// rw_hip.cpp -- random-walk Metropolis-Hastings moves written against the
// vsmch C++ layer: a Sampler whose MCMC functor moves every particle with
// vsmch::RandomWalk / RandomWalkG and NormalProposal / NormalMVProposal
// (include/vsmc/rng/random_walk.hpp), as a reference user's eval_sp would.
//
// Scenarios (the same in the oracle, oracle/rw.c):
//   0 N(0,1), NormalProposal(1.5), one-step, log gamma(x) kept
//   1 Gamma(2.5), NormalProposal(0.7, 0, inf), n-step, kept
//   2 Beta(2, 3), NormalProposal(0.9, 0, 1), one-step, not kept
//   3 -Gamma(2.5), NormalProposal(0.6, -inf, 0), n-step, not kept
//   4 3-D target, NormalMVProposal<3>(chol, a, b), n-step, kept
//   5 3-D target, block x[1..3) NormalMVProposal<2>, n-step, kept
//   6 flat, NormalProposal(1.3), one-step, not kept
//   7 flat 3-D, NormalMVProposal<3>(chol), one-step, kept
//   8 RandomWalkG, N(0,1) with g = (x, x^2), NormalProposal(1.1)
//   9 3-D target, block x[2] NormalProposal(0.5, 0, inf), one-step
//
// Usage: rw_hip scenario N iters steps seed
// Prints one line per particle: x0 x1 x2 lt accepted g0 g1 words (%.17g),
// then "ACCEPT <sum of the sampler's accept counts>".
#include <vsmch/random_walk.hpp>
#include <vsmch/sampler.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

using State = vsmch::StateHIP<8, double>;
using SP = vsmch::SingleParticleHIP<State>;

enum Slot { X0 = 0, X1 = 1, X2 = 2, LT = 3, ACC = 4, G0 = 5, G1 = 6 };

__device__ inline double lt_normal(const double *x, double *g)
{
    if (g) {
        g[0] = x[0];
        g[1] = x[0] * x[0];
    }
    return -0.5 * x[0] * x[0];
}
__device__ inline double lt_gamma(const double *x) { return 1.5 * log(x[0]) - x[0]; }
__device__ inline double lt_beta(const double *x)
{
    return log(x[0]) + 2 * log(1 - x[0]);
}
__device__ inline double lt_neg_gamma(const double *x)
{
    return 1.5 * log(-x[0]) + x[0];
}
__device__ inline double lt_mv3(const double *x)
{
    const double d = x[1] - 0.5 * x[0];
    const double l2 = x[2] > 0 ? log(x[2]) - x[2] : -INFINITY;
    return -0.5 * (x[0] * x[0] + d * d / 0.5) + l2;
}

__device__ inline int rw_dim(int s)
{
    return (s == 4 || s == 5 || s == 7 || s == 9) ? 3 : 1;
}

__device__ inline double log_target(int s, const double *x)
{
    switch (s) {
    case 0: case 8: return lt_normal(x, nullptr);
    case 1: return lt_gamma(x);
    case 2: return lt_beta(x);
    case 3: return lt_neg_gamma(x);
    case 4: case 5: case 9: return lt_mv3(x);
    default: return 0;
    }
}

__constant__ double kChol3[6] = {0.5, 0.2, 0.4, 0.1, -0.1, 0.3};
__constant__ double kA3[3] = {-INFINITY, -6, 0};
__constant__ double kB3[3] = {INFINITY, 6, INFINITY};
__constant__ double kChol2[3] = {0.4, 0.1, 0.3};
__constant__ double kA2[2] = {-6, 0};
__constant__ double kB2[2] = {6, INFINITY};

class RWInit : public vsmch::InitializeHIP<State, RWInit> {
  public:
    explicit RWInit(int s) : s_(s) {}
    __device__ std::size_t eval_sp(SP sp)
    {
        const std::size_t i = sp.id();
        double x[3] = {0, 0, 0};
        switch (s_) {
        case 0: x[0] = -2 + 4.0 * static_cast<double>(i % 97) / 96; break;
        case 1: x[0] = 0.5 + 0.25 * static_cast<double>(i % 13); break;
        case 2: x[0] = 0.05 + 0.9 * static_cast<double>(i % 17) / 16; break;
        case 3: x[0] = -0.5 - 0.3 * static_cast<double>(i % 9); break;
        case 4: case 5: case 9:
            x[0] = 0.1 * static_cast<double>(i % 5) - 0.2;
            x[1] = 0.5;
            x[2] = 1 + 0.1 * static_cast<double>(i % 7);
            break;
        case 7: x[0] = 1; x[1] = 2; x[2] = 3; break;
        case 8: x[0] = 0.3 * static_cast<double>(i % 7); break;
        default: break;
        }
        for (int k = 0; k != 3; ++k)
            sp.state(X0 + k) = x[k];
        sp.state(LT) = log_target(s_, x);
        sp.state(ACC) = 0;
        sp.state(G0) = s_ == 8 ? x[0] : 0;
        sp.state(G1) = s_ == 8 ? x[0] * x[0] : 0;
        return 0;
    }

  private:
    int s_;
};

class RWMove : public vsmch::MoveHIP<State, RWMove> {
  public:
    RWMove(int s, std::size_t steps) : s_(s), steps_(steps) {}

    __device__ std::size_t eval_sp(std::size_t, SP sp)
    {
        double x[3] = {sp.state(X0), sp.state(X1), sp.state(X2)};
        double lt = sp.state(LT);
        double g[2] = {sp.state(G0), sp.state(G1)};
        auto &rng = sp.rng();
        const std::size_t n = steps_;
        std::size_t acc = 0;
        auto lt1 = [this](std::size_t, const double *v) { return log_target(s_, v); };
        switch (s_) {
        case 0: { // one-step, ltx kept
            vsmch::RandomWalk<double, 1> rw;
            vsmch::NormalProposal<double> prop(1.5);
            for (std::size_t k = 0; k != n; ++k)
                acc += rw(rng, x, &lt, lt1, prop);
            break;
        }
        case 1: { // n-step, ltx kept
            vsmch::RandomWalk<double, 1> rw;
            vsmch::NormalProposal<double> prop(0.7, 0, INFINITY);
            acc += rw(n, rng, x, &lt, lt1, prop);
            break;
        }
        case 2: { // one-step, ltx null
            vsmch::RandomWalk<double, 1> rw;
            vsmch::NormalProposal<double> prop(0.9, 0, 1);
            for (std::size_t k = 0; k != n; ++k)
                acc += rw(rng, x, static_cast<double *>(nullptr), lt1, prop);
            lt = log_target(s_, x);
            break;
        }
        case 3: { // n-step, ltx null
            vsmch::RandomWalk<double, 1> rw;
            vsmch::NormalProposal<double> prop(0.6, -INFINITY, 0);
            acc += rw(n, rng, x, static_cast<double *>(nullptr), lt1, prop);
            lt = log_target(s_, x);
            break;
        }
        case 4: { // 3-D, n-step, ltx kept
            vsmch::RandomWalk<double, 3> rw;
            vsmch::NormalMVProposal<double, 3> prop(kChol3, kA3, kB3);
            acc += rw(n, rng, x, &lt, lt1, prop);
            break;
        }
        case 5: { // block x[1..3), n-step, ltx kept
            vsmch::RandomWalk<double, 2> rw;
            vsmch::NormalMVProposal<double, 2> prop(kChol2, kA2, kB2);
            acc += rw(n, rng, 3, 1, x, &lt, lt1, prop);
            break;
        }
        case 6: { // flat, one-step, ltx null
            vsmch::RandomWalk<double, 1> rw;
            vsmch::NormalProposal<double> prop(1.3);
            for (std::size_t k = 0; k != n; ++k)
                acc += rw(rng, x, static_cast<double *>(nullptr), lt1, prop);
            lt = log_target(s_, x);
            break;
        }
        case 7: { // flat 3-D, one-step, ltx kept
            vsmch::RandomWalk<double, 3> rw;
            vsmch::NormalMVProposal<double, 3> prop(kChol3);
            for (std::size_t k = 0; k != n; ++k)
                acc += rw(rng, x, &lt, lt1, prop);
            break;
        }
        case 8: { // RandomWalkG, one-step, ltx and g kept
            vsmch::RandomWalkG<double, 1, 2> rw;
            vsmch::NormalProposal<double> prop(1.1);
            auto ltg = [](std::size_t, std::size_t, const double *v, double *gv) {
                return lt_normal(v, gv);
            };
            for (std::size_t k = 0; k != n; ++k)
                acc += rw(rng, x, &lt, g, ltg, prop);
            break;
        }
        case 9: { // block x[2], one-step, ltx null
            vsmch::RandomWalk<double, 1> rw;
            vsmch::NormalProposal<double> prop(0.5, 0, INFINITY);
            for (std::size_t k = 0; k != n; ++k)
                acc += rw(rng, 3, 2, x, static_cast<double *>(nullptr), lt1, prop);
            lt = log_target(s_, x);
            break;
        }
        default: break;
        }
        sp.state(X0) = x[0];
        sp.state(X1) = x[1];
        sp.state(X2) = x[2];
        sp.state(LT) = lt;
        sp.state(ACC) += static_cast<double>(acc);
        sp.state(G0) = g[0];
        sp.state(G1) = g[1];
        return acc;
    }

  private:
    int s_;
    std::size_t steps_;
};

int main(int argc, char **argv)
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: rw_hip scenario N iters steps seed\n");
        return 2;
    }
    const int s = std::atoi(argv[1]);
    const std::size_t N = std::strtoull(argv[2], nullptr, 10);
    const std::size_t iters = std::strtoull(argv[3], nullptr, 10);
    const std::size_t steps = std::strtoull(argv[4], nullptr, 10);
    const unsigned seed = static_cast<unsigned>(std::strtoul(argv[5], nullptr, 10));
    if (s < 0 || s > 9 || N == 0 || iters == 0) {
        std::fprintf(stderr, "rw_hip: scenario 0..9, N > 0, iters > 0\n");
        return 2;
    }
    try {
        vsmch::Seed::instance().set(seed);
        vsmch::Sampler<State> sampler(N, vsmch::Stratified,
            vsmch::Sampler<State>::resample_threshold_never());
        sampler.init(RWInit(s)).mcmc(RWMove(s, steps), false);
        sampler.initialize().iterate(iters - 1);
        std::size_t acc_total = 0;
        for (std::size_t t = 0; t != sampler.iter_size(); ++t)
            for (std::size_t k = 0; k != 2; ++k)
                acc_total += sampler.accept_history(k, t);
        std::vector<double> col[7];
        for (int k = 0; k != 7; ++k) {
            col[k].resize(N);
            sampler.particle().value().read_state(k, col[k].data());
        }
        std::vector<unsigned long long> words(N);
        vsmch::check_hip(hipMemcpy(words.data(), sampler.particle().rng_words(),
                             sizeof(unsigned long long) * N, hipMemcpyDeviceToHost),
            "words");
        for (std::size_t i = 0; i != N; ++i) {
            for (int k = 0; k != 7; ++k)
                std::printf("%.17g ", col[k][i]);
            std::printf("%llu\n", words[i]);
        }
        std::printf("ACCEPT %zu\n", acc_total);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "rw_hip: %s\n", e.what());
        return 1;
    }
    return 0;
}
