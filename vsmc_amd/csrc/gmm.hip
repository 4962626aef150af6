// gmm.hip -- tempered SMC sampler for a Gaussian mixture on the device.
//
// The model, prior, moves and proposal scales are those of
// example/gmm/include/gmm.hpp (gmm_state, gmm_init, gmm_move_smc,
// gmm_move_mu / _lambda / _weight, gmm_path_integrand); the tempering is
// adaptive: each step picks the largest increment of alpha whose
// conditional ESS stays at the target (the reference anneals on a fixed
// schedule, gmm.hpp:567-604). Per step: search for the increment
// (multi-candidate bisection on the device), Weight::add_log of
// increment x log-likelihood (log-evidence accumulated), stratified
// resampling when ESS < threshold N (the particle set's exact path, all
// rows gathered), then `moves` random-walk Metropolis-Hastings moves that
// cycle mu, lambda, weight, and the path-sampling monitor
// sum_i W_i log-likelihood_i.
//
// A particle is 3c + 2 FP32 rows of a particle set (component stride N):
// mu[c], lambda[c], weight[c], log prior, log likelihood. The data sit in
// LDS; every likelihood evaluation is c exponentials and one logarithm per
// data point (FP32, hardware exp2 / log2).
#include "particle.hpp"
#include "../../include/vsmch/device_rng.hpp"

#include <math.h>
#include <type_traits>
#include <vector>

namespace vsmch {

constexpr int kGmmMaxC = 8;
constexpr int kGmmMaxData = 8192;
constexpr int kGB = 256;           // threads per block
constexpr int kCand = 16;          // candidates per search round

struct GmmDev {
    double alpha;
    double dalpha;
    double lo, hi;          // search bracket of the increment
    double llmax;
    double mu0, sd0, shape0, scale0;
    double target;          // conditional ESS target (fraction)
    unsigned long long accepted[3];
    unsigned long long proposed[3];
};

struct GmmSet {
    ParticleSet *p = nullptr;
    GmmDev *g = nullptr;
    float *data = nullptr;
    int ndata = 0;
    int c = 4;
    int moves = 5;
    int rounds = 8;
    int scheme = kStratified;
    double threshold = 0.5;
    double *part = nullptr;   // search partials: 2 kCand per block
    double *inc = nullptr;    // per-particle increments (n)
    long long nb = 0;
};

namespace {

__device__ inline float gmm_log_prior(int c, const float *mu, const float *lam,
    const GmmDev &g)
{
    const float mu0 = static_cast<float>(g.mu0);
    const float iv = static_cast<float>(1 / (g.sd0 * g.sd0));
    const float a1 = static_cast<float>(g.shape0 - 1);
    const float is = static_cast<float>(1 / g.scale0);
    float lp = 0.0f;
    for (int i = 0; i < c; ++i) {
        const float r = mu[i] - mu0;
        lp += -0.5f * r * r * iv + a1 * logf(lam[i]) - lam[i] * is;
    }
    return lp;
}

// sum_k log(sum_i w_i exp(log(lambda_i) / 2 - lambda_i (x_k - mu_i)^2 / 2)
// + 1e-13) - n log(2 pi) / 2, the data in LDS. The component count is a
// compile-time constant so that exactly C exponentials are issued per point.
// Each exponent, in log2 units, is a_i - r^2 with r = s_i x + m_i,
// s_i = sqrt(lambda_i log2(e) / 2), m_i = -mu_i s_i: two FMAs, then the
// hardware exp2. The logarithm is taken of the product of two consecutive
// points' mixture densities (each at least 1e-13 and at most
// max_i sqrt(lambda_i), so the product stays in the normal FP32 range) and
// accumulated in log2.
template <int C>
__device__ inline float gmm_mix(float x, const float *a, const float *sc,
    const float *m)
{
    float s = 1e-13f;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        const float r = fmaf(x, sc[i], m[i]);
        s += __builtin_amdgcn_exp2f(fmaf(-r, r, a[i]));
    }
    return s;
}

template <int C>
__device__ inline float gmm_log_lik(const float *mu, const float *lam,
    const float *w, const float *xs, int n)
{
    float a[C], sc[C], m[C];
    const float l2e = 1.4426950408889634f;
#pragma unroll
    for (int i = 0; i < C; ++i) {
        a[i] = (__logf(w[i]) + 0.5f * __logf(lam[i])) * l2e;
        sc[i] = sqrtf(0.5f * l2e * lam[i]);
        m[i] = -mu[i] * sc[i];
    }
    float ll2 = 0.0f;
    int k = 0;
    for (; k + 1 < n; k += 2) {
        const float s0 = gmm_mix<C>(xs[k], a, sc, m);
        const float s1 = gmm_mix<C>(xs[k + 1], a, sc, m);
        ll2 += __builtin_amdgcn_logf(s0 * s1);
    }
    if (k < n)
        ll2 += __builtin_amdgcn_logf(gmm_mix<C>(xs[k], a, sc, m));
    return ll2 * 0.69314718055994531f -
        0.91893853320467274178f * static_cast<float>(n);
}

__device__ inline float stream_normal(ParticleStream &s, float &cache,
    bool &have)
{
    if (have) {
        have = false;
        return cache;
    }
    const uint32_t a = s.next();
    const uint32_t b = s.next();
    const normal_pair<float> p = box_muller<float>(a, b);
    cache = p.z1;
    have = true;
    return p.z0;
}

// Gamma(shape >= 1, scale) by Marsaglia and Tsang (2000). Each attempt is
// accepted with probability at least 0.95 for shape >= 1 (0.98 at the prior's
// shape 2), so 32 straight rejections have probability below 0.05^32 = 2e-42
// per draw; the bound is there so that a NaN or infinite shape cannot spin,
// and its fallback (the mean) is never reached by a finite shape in practice.
__device__ inline float stream_gamma(ParticleStream &s, float shape,
    float scale, float &cache, bool &have)
{
    const float d = shape - 1.0f / 3.0f;
    const float cc = 1.0f / sqrtf(9.0f * d);
    for (int it = 0; it < 32; ++it) {
        const float z = stream_normal(s, cache, have);
        float v = 1.0f + cc * z;
        if (v <= 0.0f)
            continue;
        v = v * v * v;
        const float u = u01_u32f(s.next());
        if (logf(u) < 0.5f * z * z + d - d * v + d * logf(v))
            return d * v * scale;
    }
    return shape * scale;
}

__device__ inline void load_data(float *xs, const float *data, int n)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x)
        xs[k] = data[k];
    __syncthreads();
}

// Initialisation (gmm_init::eval_sp, gmm.hpp:364-387): mu ~ N(mu0, sd0),
// lambda ~ Gamma(shape0, scale0), weights Gamma(1, 1) normalised; words
// [0, 64 (3c + 1)) of particle i's stream.
template <int C>
__global__ void __launch_bounds__(kGB) k_gmm_init(float *state, long long n,
    uint32_t seed, const GmmDev *gp, const float *data, int ndata)
{
    constexpr int c = C;
    extern __shared__ float xs[];
    load_data(xs, data, ndata);
    const long long i = static_cast<long long>(blockIdx.x) * kGB + threadIdx.x;
    if (i >= n)
        return;
    const GmmDev g = *gp;
    ParticleStream s(seed_key(seed, static_cast<unsigned long long>(i)), 0);
    float cache = 0.0f;
    bool have = false;
    float mu[C], lam[C], w[C];
    float sum = 0.0f;
    for (int k = 0; k < c; ++k) {
        mu[k] = static_cast<float>(g.mu0) + static_cast<float>(g.sd0) *
            stream_normal(s, cache, have);
        lam[k] = stream_gamma(s, static_cast<float>(g.shape0),
            static_cast<float>(g.scale0), cache, have);
        w[k] = -logf(1.0f - u01_u32f(s.next()));
        sum += w[k];
    }
    for (int k = 0; k < c; ++k)
        w[k] /= sum;
    for (int k = 0; k < c; ++k) {
        state[k * n + i] = mu[k];
        state[(c + k) * n + i] = lam[k];
        state[(2 * c + k) * n + i] = w[k];
    }
    state[3 * c * n + i] = gmm_log_prior(c, mu, lam, g);
    state[(3 * c + 1) * n + i] = gmm_log_lik<C>(mu, lam, w, xs, ndata);
}

__device__ inline float logit_side(int c, const float *w)
{
    float sum_lw = 1.0f, sum_llw = 0.0f;
    for (int i = 0; i < c - 1; ++i) {
        const float r = w[i] / w[c - 1];
        sum_lw += r;
        sum_llw += logf(r);
    }
    return sum_llw - static_cast<float>(c) * logf(sum_lw);
}

// One Metropolis-Hastings move of `kind` (0 mu, 1 lambda, 2 weight) at
// temperature alpha (gmm.hpp:447-543). Draws: c normals then one uniform
// from word 64 (32 t + m + 1) of particle i's stream.
template <int C>
__global__ void __launch_bounds__(kGB) k_gmm_move(float *state, long long n,
    int kind, uint32_t seed, long long t, int m, GmmDev *gp,
    const float *data, int ndata)
{
    constexpr int c = C;
    extern __shared__ float xs[];
    load_data(xs, data, ndata);
    __shared__ long long acc_lds[kGB / 64];
    const long long i = static_cast<long long>(blockIdx.x) * kGB + threadIdx.x;
    int acc = 0;
    if (i < n) {
        const GmmDev &g = *gp;
        const double alpha_d = g.alpha;
        const float alpha = static_cast<float>(alpha_d);
        const double a = alpha_d < 0.02 ? 0.02 : alpha_d;
        const float sd_mu = static_cast<float>(0.15 / a);
        const float sd_lam = static_cast<float>((1 + sqrt(1 / a)) * 0.15);
        const float sd_w = static_cast<float>((1 + sqrt(1 / a)) * 0.2);
        float mu[C], lam[C], w[C];
        for (int k = 0; k < c; ++k) {
            mu[k] = state[k * n + i];
            lam[k] = state[(c + k) * n + i];
            w[k] = state[(2 * c + k) * n + i];
        }
        const float lp0 = state[3 * c * n + i];
        const float ll0 = state[(3 * c + 1) * n + i];
        ParticleStream s(seed_key(seed, static_cast<unsigned long long>(i)),
            64ULL * static_cast<unsigned long long>(32 * t + m + 1));
        float cache = 0.0f;
        bool have = false;
        float z[C];
        for (int k = 0; k < c; ++k)
            z[k] = stream_normal(s, cache, have);
        const float log_u = logf(u01_u32f(s.next()));
        float jac = 0.0f;
        float w2[C], mu2[C], lam2[C];
        for (int k = 0; k < c; ++k) {
            mu2[k] = mu[k];
            lam2[k] = lam[k];
            w2[k] = w[k];
        }
        if (kind == 0) {
            for (int k = 0; k < c; ++k)
                mu2[k] = mu[k] + sd_mu * z[k];
        } else if (kind == 1) {
            for (int k = 0; k < c; ++k) {
                lam2[k] = lam[k] * expf(sd_lam * z[k]);
                jac += logf(lam2[k] / lam[k]);
            }
        } else {
            float sum = 1.0f;
            for (int k = 0; k < c - 1; ++k) {
                w2[k] = expf(logf(w[k] / w[c - 1]) + sd_w * z[k]);
                sum += w2[k];
            }
            w2[c - 1] = 1.0f;
            for (int k = 0; k < c; ++k)
                w2[k] /= sum;
            jac = logit_side(c, w2) - logit_side(c, w);
        }
        const float lp1 = gmm_log_prior(c, mu2, lam2, g);
        const float ll1 = gmm_log_lik<C>(mu2, lam2, w2, xs, ndata);
        const float pr = jac + lp1 + alpha * ll1 - (lp0 + alpha * ll0);
        // The reference rejects when pr < log u (mh_reject_common). A float
        // uniform of exactly 0 (probability 2^-24) makes an infinite normal
        // and can make pr NaN (inf - inf); such a proposal is rejected here
        // instead of accepted.
        if (pr >= log_u) {
            acc = 1;
            for (int k = 0; k < c; ++k) {
                state[k * n + i] = mu2[k];
                state[(c + k) * n + i] = lam2[k];
                state[(2 * c + k) * n + i] = w2[k];
            }
            state[3 * c * n + i] = lp1;
            state[(3 * c + 1) * n + i] = ll1;
        }
    }
    const long long a = block_reduce(static_cast<long long>(acc), OpAddL(),
        acc_lds);
    if (threadIdx.x == 0) {
        atomicAdd(&gp->accepted[kind], static_cast<unsigned long long>(a));
        const long long b0 = static_cast<long long>(blockIdx.x) * kGB;
        atomicAdd(&gp->proposed[kind], static_cast<unsigned long long>(
            n - b0 < kGB ? n - b0 : kGB));
    }
}

// Per-block maximum of the log-likelihood row.
__global__ void __launch_bounds__(kGB) k_gmm_llmax(const float *ll,
    long long n, double *part)
{
    __shared__ double lds[kGB / 64];
    double m = -INFINITY;
    for (long long i = static_cast<long long>(blockIdx.x) * kGB + threadIdx.x;
         i < n; i += static_cast<long long>(gridDim.x) * kGB) {
        const double v = ll[i];
        m = v > m ? v : m;
    }
    m = block_reduce(m, OpMaxD(), lds);
    if (threadIdx.x == 0)
        part[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kGB) k_gmm_search_start(const double *part,
    long long nb, GmmDev *g)
{
    __shared__ double lds[kGB / 64];
    double m = -INFINITY;
    for (long long b = threadIdx.x; b < nb; b += kGB)
        m = part[b] > m ? part[b] : m;
    m = block_reduce(m, OpMaxD(), lds);
    if (threadIdx.x == 0) {
        g->llmax = m;
        g->lo = 0.0;
        g->hi = 1.0 - g->alpha;
    }
}

__device__ inline double cand(const GmmDev &g, int k)
{
    return g.lo + (g.hi - g.lo) * static_cast<double>(k + 1) / kCand;
}

// Per-block sums S1_k = sum W_i e^{d_k x_i}, S2_k = sum W_i e^{2 d_k x_i},
// x_i = l_i - lmax <= 0, for the kCand equally spaced candidates
// d_k = lo + (k + 1) h of the bracket, h = (hi - lo) / kCand. Two FP64
// exponentials per particle: e^{d_0 x}, then k steps of the ratio e^{h x}
// (relative error of a term about 1e-13 at worst, far below the bracket's
// resolution). Written as part[block][0 .. kCand) = S1, [kCand, 2 kCand) = S2.
__global__ void __launch_bounds__(kGB) k_gmm_search_pass(const float *ll,
    const double *logw, const PDev *pd, long long n, const GmmDev *gp,
    double *part)
{
    __shared__ double lds[kGB / kWave][2 * kCand];
    const GmmDev g = *gp;
    const double h = (g.hi - g.lo) / kCand;
    const double d0 = cand(g, 0);
    double s1[kCand], s2[kCand];
#pragma unroll
    for (int k = 0; k < kCand; ++k)
        s1[k] = s2[k] = 0.0;
    const int equal = pd->equal;
    const double c = pd->c;
    const double w_equal = 1.0 / static_cast<double>(n);
    for (long long i = static_cast<long long>(blockIdx.x) * kGB + threadIdx.x;
         i < n; i += static_cast<long long>(gridDim.x) * kGB) {
        // normalised weights: the conditional ESS is a fraction of 1
        const double W = equal ? w_equal : exp(logw[i] - c);
        const double x = static_cast<double>(ll[i]) - g.llmax;
        const double r = exp(h * x);
        double e = exp(d0 * x);
#pragma unroll
        for (int k = 0; k < kCand; ++k) {
            const double we = W * e;
            s1[k] += we;
            s2[k] = fma(we, e, s2[k]);
            e *= r;
        }
    }
    // s1 and s2 as one array of 2 kCand values; lane l then holds the
    // wave's total of value l / 2
    double v[2 * kCand];
#pragma unroll
    for (int k = 0; k < kCand; ++k) {
        v[k] = s1[k];
        v[kCand + k] = s2[k];
    }
    const double tot = wave_transpose_sum<2 * kCand>(v);
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    static_assert(kWave == 4 * kCand, "two lanes per value");
    if ((lane & 1) == 0)
        lds[wave][lane >> 1] = tot;
    __syncthreads();
    if (threadIdx.x < 2 * kCand) {
        double t = lds[0][threadIdx.x];
        for (int w = 1; w < kGB / kWave; ++w)
            t += lds[w][threadIdx.x];
        part[blockIdx.x * (2 * kCand) + threadIdx.x] = t;
    }
}

// Sums the blocks' partials (in a fixed order: deterministic), then narrows
// the bracket to [d_j, d_{j+1}] around the last candidate j whose
// conditional ESS S1^2 / S2 reaches the target (d_{-1} = lo); final:
// dalpha = lo, or hi when no candidate reached it (a step smaller than the
// resolution).
__global__ void __launch_bounds__(kGB) k_gmm_search_merge(const double *part,
    long long nb, GmmDev *g, int final_round)
{
    constexpr int kV = 2 * kCand;          // values per block
    constexpr int kSlices = kGB / kV;      // threads per value
    __shared__ double lds[kSlices][kV];
    __shared__ double cs[kCand];
    const int v = threadIdx.x % kV;
    const int sl = threadIdx.x / kV;
    // four independent chains so that the loads overlap
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    long long j = sl;
    for (; j + 3 * kSlices < nb; j += 4 * kSlices) {
        a0 += part[j * kV + v];
        a1 += part[(j + kSlices) * kV + v];
        a2 += part[(j + 2 * kSlices) * kV + v];
        a3 += part[(j + 3 * kSlices) * kV + v];
    }
    for (; j < nb; j += kSlices)
        a0 += part[j * kV + v];
    lds[sl][v] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (threadIdx.x < kCand) {
        double s1 = 0.0, s2 = 0.0;
        for (int q = 0; q < kSlices; ++q) {
            s1 += lds[q][threadIdx.x];
            s2 += lds[q][kCand + threadIdx.x];
        }
        cs[threadIdx.x] = s1 * s1 / s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const GmmDev hd = *g;
        int j = -1;
        for (int k = 0; k < kCand; ++k)
            if (cs[k] >= hd.target)
                j = k;
        const double lo = j < 0 ? hd.lo : cand(hd, j);
        const double hi = j + 1 < kCand ? cand(hd, j + 1) : cand(hd, kCand - 1);
        g->lo = lo;
        g->hi = j == kCand - 1 ? lo : hi;
        if (final_round)
            g->dalpha = lo > 0.0 ? lo : hi;
    }
}

// Weight increments dalpha * log-likelihood (FP64) for Weight::add_log.
__global__ void __launch_bounds__(kGB) k_gmm_increment(const float *ll,
    long long n, const GmmDev *g, double *inc)
{
    const long long i = static_cast<long long>(blockIdx.x) * kGB + threadIdx.x;
    if (i < n)
        inc[i] = g->dalpha * static_cast<double>(ll[i]);
}

__global__ void k_gmm_advance(GmmDev *g, double *alpha_hist, long long t)
{
    double a = g->alpha + g->dalpha;
    g->alpha = a < 1.0 ? a : 1.0;
    alpha_hist[t] = g->alpha;
    for (int k = 0; k < 3; ++k) {
        g->accepted[k] = 0;
        g->proposed[k] = 0;
    }
}

__global__ void k_gmm_store(const PDev *pd, const GmmDev *g, long long t,
    double *logz_hist, double *acc_hist)
{
    logz_hist[t] = pd->logz;
    for (int k = 0; k < 3; ++k)
        acc_hist[3 * t + k] = g->proposed[k] ?
            static_cast<double>(g->accepted[k]) / static_cast<double>(g->proposed[k]) :
            -1.0;
}

size_t data_bytes(const GmmSet *s)
{
    return sizeof(float) * static_cast<size_t>(s->ndata);
}

} // namespace

void gmm_destroy(GmmSet *s)
{
    if (!s)
        return;
    if (s->p)
        particle_destroy(s->p);
    (void)hipFree(s->g);
    (void)hipFree(s->data);
    (void)hipFree(s->part);
    (void)hipFree(s->inc);
    delete s;
}

int gmm_create(GmmSet **out, long long n, int c, uint32_t seed, double cess,
    int moves, int rounds, int scheme, double threshold, const double *data,
    long long ndata, hipStream_t st)
{
    if (n < 2 || c < 2 || c > kGmmMaxC || ndata < 2 || ndata > kGmmMaxData ||
        !(cess > 0 && cess < 1) || moves < 0 || rounds < 1 ||
        (scheme != kStratified && scheme != kSystematic)) {
        set_error("gmm_create: need n >= 2, 2 <= c <= 8, 2 <= data <= 8192, "
                  "0 < cess < 1, stratified or systematic");
        return kErrInvalid;
    }
    GmmSet *s = new GmmSet();
    s->c = c;
    s->moves = moves;
    s->rounds = rounds;
    s->scheme = scheme;
    s->threshold = threshold;
    s->ndata = static_cast<int>(ndata);
    int rc = particle_create(&s->p, n, 3 * c + 2, 4, seed, st);
    if (rc != kOk) {
        delete s;
        return rc;
    }
    // search blocks: two per CU, each thread looping over particles
    s->nb = (n + kGB - 1) / kGB;
    if (s->nb > 512)
        s->nb = 512;
    std::vector<float> xf(ndata);
    for (long long k = 0; k < ndata; ++k)
        xf[k] = static_cast<float>(data[k]);
    double lo = data[0], hi = data[0];
    for (long long k = 0; k < ndata; ++k) {
        lo = data[k] < lo ? data[k] : lo;
        hi = data[k] > hi ? data[k] : hi;
    }
    GmmDev g{};
    const double kappa = 1 / ((hi - lo) * (hi - lo));
    g.mu0 = 0.5 * (lo + hi);
    g.sd0 = 1 / std::sqrt(kappa);
    g.shape0 = 2;
    g.scale0 = 50 * kappa;
    g.target = cess;
    if (hipMalloc(&s->g, sizeof(GmmDev)) != hipSuccess ||
        hipMalloc(&s->data, sizeof(float) * ndata) != hipSuccess ||
        hipMalloc(&s->part, sizeof(double) * 2 * kCand * s->nb) != hipSuccess ||
        hipMalloc(&s->inc, sizeof(double) * n) != hipSuccess ||
        hipMemcpy(s->g, &g, sizeof(g), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->data, xf.data(), sizeof(float) * ndata,
            hipMemcpyHostToDevice) != hipSuccess) {
        gmm_destroy(s);
        set_error("gmm_create: device allocation failed");
        return kErrAlloc;
    }
    *out = s;
    return kOk;
}

ParticleSet *gmm_particles(GmmSet *s) { return s->p; }

// Calls f(std::integral_constant<int, c>) for the component counts the
// kernels are instantiated for (2 .. kGmmMaxC, checked by gmm_create).
template <typename F>
int with_components(int c, F &&f)
{
    switch (c) {
    case 2: return f(std::integral_constant<int, 2>());
    case 3: return f(std::integral_constant<int, 3>());
    case 4: return f(std::integral_constant<int, 4>());
    case 5: return f(std::integral_constant<int, 5>());
    case 6: return f(std::integral_constant<int, 6>());
    case 7: return f(std::integral_constant<int, 7>());
    case 8: return f(std::integral_constant<int, 8>());
    default:
        set_error("gmm: component count outside 2 .. 8");
        return kErrInvalid;
    }
}

// One move of `kind` over all particles at the current alpha (tests).
int gmm_move(GmmSet *s, int kind, long long t, int m)
{
    ParticleSet *p = s->p;
    const unsigned g = static_cast<unsigned>((p->n + kGB - 1) / kGB);
    return with_components(s->c, [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        hipLaunchKernelGGL(k_gmm_move<C>, dim3(g), dim3(kGB), data_bytes(s),
            p->ws.stream, static_cast<float *>(p->state), p->n, kind, p->seed,
            t, m, s->g, static_cast<const float *>(s->data), s->ndata);
        VSMCH_HIP(hipGetLastError());
        return kOk;
    });
}

int gmm_set_alpha(GmmSet *s, double alpha)
{
    VSMCH_HIP(hipMemcpyAsync(&s->g->alpha, &alpha, sizeof(double),
        hipMemcpyHostToDevice, s->p->ws.stream));
    VSMCH_HIP(hipStreamSynchronize(s->p->ws.stream));
    return kOk;
}

// The whole sampler: initialise at alpha = 0, then steps until alpha = 1
// (at most max_steps entries of history, the initial one included).
// Histories (device): alpha, ESS before resampling, resampled flags,
// cumulative log-evidence, the path-sampling integrand, acceptance rates
// [3 per step]. steps_out: entries written.
int gmm_run(GmmSet *s, long long max_steps, double *alpha_d, double *ess_d,
    int *rs_d, double *logz_d, double *integ_d, double *acc_d,
    long long *steps_out)
{
    ParticleSet *p = s->p;
    hipStream_t st = p->ws.stream;
    const long long n = p->n;
    const int c = s->c;
    const unsigned gp = static_cast<unsigned>((n + kGB - 1) / kGB);
    float *state = static_cast<float *>(p->state);
    const float *ll = state + (3 * c + 1) * n;
    GmmDev g0{};
    VSMCH_HIP(hipMemcpy(&g0, s->g, sizeof(g0), hipMemcpyDeviceToHost));
    g0.alpha = 0;
    g0.dalpha = 0;
    VSMCH_HIP(hipMemcpy(s->g, &g0, sizeof(g0), hipMemcpyHostToDevice));
    int rc = with_components(c, [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        hipLaunchKernelGGL(k_gmm_init<C>, dim3(gp), dim3(kGB), data_bytes(s),
            st, state, n, p->seed, static_cast<const GmmDev *>(s->g),
            static_cast<const float *>(s->data), s->ndata);
        VSMCH_HIP(hipGetLastError());
        return kOk;
    });
    if (rc != kOk)
        return rc;
    rc = weight_update(p, nullptr, 2);   // set_equal
    if (rc != kOk)
        return rc;
    // step 0 records
    const double zero = 0.0, nd = static_cast<double>(n);
    const int izero = 0;
    VSMCH_HIP(hipMemcpyAsync(alpha_d, &zero, sizeof(double), hipMemcpyHostToDevice, st));
    VSMCH_HIP(hipMemcpyAsync(ess_d, &nd, sizeof(double), hipMemcpyHostToDevice, st));
    VSMCH_HIP(hipMemcpyAsync(rs_d, &izero, sizeof(int), hipMemcpyHostToDevice, st));
    VSMCH_HIP(hipMemcpyAsync(logz_d, &zero, sizeof(double), hipMemcpyHostToDevice, st));
    rc = moments(p, 3 * c + 1, 1, integ_d);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL(k_gmm_store, dim3(1), dim3(1), 0, st,
        static_cast<const PDev *>(p->pd), static_cast<const GmmDev *>(s->g),
        0LL, logz_d, acc_d);
    long long t = 1;
    double alpha = 0;
    for (; t < max_steps && alpha < 1.0; ++t) {
        // increment of alpha with conditional ESS = target
        hipLaunchKernelGGL(k_gmm_llmax, dim3(s->nb), dim3(kGB), 0, st, ll, n,
            s->part);
        hipLaunchKernelGGL(k_gmm_search_start, dim3(1), dim3(kGB), 0, st,
            static_cast<const double *>(s->part), s->nb, s->g);
        for (int r = 0; r < s->rounds; ++r) {
            hipLaunchKernelGGL(k_gmm_search_pass, dim3(s->nb), dim3(kGB), 0,
                st, ll, static_cast<const double *>(p->logw),
                static_cast<const PDev *>(p->pd), n,
                static_cast<const GmmDev *>(s->g), s->part);
            hipLaunchKernelGGL(k_gmm_search_merge, dim3(1), dim3(kGB), 0, st,
                static_cast<const double *>(s->part), s->nb, s->g,
                r == s->rounds - 1 ? 1 : 0);
        }
        // reweight (gmm_move_smc: add_log of increment x log-likelihood)
        hipLaunchKernelGGL(k_gmm_increment, dim3(gp), dim3(kGB), 0, st, ll, n,
            static_cast<const GmmDev *>(s->g), s->inc);
        VSMCH_HIP(hipGetLastError());
        rc = weight_update(p, s->inc, 1);
        if (rc != kOk)
            return rc;
        hipLaunchKernelGGL(k_gmm_advance, dim3(1), dim3(1), 0, st, s->g,
            alpha_d, t);
        rc = resample_step(p, s->scheme, s->threshold, t, ess_d, rs_d, nullptr,
            nullptr);
        if (rc != kOk)
            return rc;
        for (int m = 0; m < s->moves; ++m) {
            rc = gmm_move(s, m % 3, t, m);
            if (rc != kOk)
                return rc;
        }
        rc = moments(p, 3 * c + 1, 1, integ_d + t);
        if (rc != kOk)
            return rc;
        hipLaunchKernelGGL(k_gmm_store, dim3(1), dim3(1), 0, st,
            static_cast<const PDev *>(p->pd), static_cast<const GmmDev *>(s->g),
            t, logz_d, acc_d);
        VSMCH_HIP(hipGetLastError());
        VSMCH_HIP(hipMemcpyAsync(&alpha, &s->g->alpha, sizeof(double),
            hipMemcpyDeviceToHost, st));
        VSMCH_HIP(hipStreamSynchronize(st));
    }
    *steps_out = t;
    return kOk;
}

} // namespace vsmch
