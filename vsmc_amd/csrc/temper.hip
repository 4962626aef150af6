// temper.hip -- conditional ESS of a tempering increment over a particle
// set's weights, and the largest increment that keeps it at a target: the
// adaptive step of tempered SMC for any target whose per-particle log
// likelihoods the caller holds on the device. (The reference anneals on a
// fixed schedule, example/gmm/include/gmm.hpp:567-604; the same search runs
// inside the mixture sampler.)
//
// CESS(d) = (sum_i W_i e^{d l_i})^2 / sum_i W_i e^{2 d l_i}, W the current
// normalised weights, in (0, 1] and non-increasing in d >= 0. Terms are
// taken relative to m = max_i l_i (the ratio does not change); a particle
// whose l_i is -inf contributes nothing for d > 0 (e^{0 l_i} = 1 at d = 0),
// one whose l_i is NaN nothing at all.
#include "particle.hpp"

namespace vsmch {

namespace {
constexpr int kPB = 256;                 // threads per block
constexpr int kPItems = 16;              // particles per thread
constexpr int kPTile = kPB * kPItems;    // particles per block
constexpr int kCand = 16;                // candidates per search round
} // namespace

// Search state (device): bracket [lo, hi] with CESS(lo) >= target
// (CESS(0) = 1) and the candidates of the next pass.
struct CessDev {
    double m;              // max_i l_i
    double lo, hi;
    double target;
    double cand[kCand];
    int ncand;
    int done;              // CESS(hi) >= target: the answer is hi
    double cess[kCand];    // CESS of the last pass's candidates
};

__global__ void __launch_bounds__(kPB) k_ll_max(const double *ll, long long n,
    double *part)
{
    const long long base = static_cast<long long>(blockIdx.x) * kPTile;
    double m = -INFINITY;
    for (int j = 0; j < kPItems; ++j) {
        const long long i = base + j * kPB + threadIdx.x;
        if (i < n) {
            const double v = ll[i];
            m = v > m ? v : m;   // NaN compares false: skipped
        }
    }
    __shared__ double lds[kWaves];
    m = block_reduce(m, OpMaxD(), lds);
    if (threadIdx.x == 0)
        part[blockIdx.x] = m;
}

// The maximum over the block maxima, and the first pass's candidates: the
// single increment d (solve = 0) or dmax itself (solve = 1, checked first).
__global__ void __launch_bounds__(kPB) k_cess_start(const double *part,
    long long nb, double d, double target, int solve, CessDev *cd)
{
    double m = -INFINITY;
    for (long long b = threadIdx.x; b < nb; b += kPB) {
        const double v = part[b];
        m = v > m ? v : m;
    }
    __shared__ double lds[kWaves];
    m = block_reduce(m, OpMaxD(), lds);
    if (threadIdx.x == 0) {
        cd->m = m;
        cd->lo = 0.0;
        cd->hi = d;
        cd->target = target;
        cd->cand[0] = d;
        cd->ncand = 1;
        cd->done = 0;
    }
}

// Per-block sums s1_k = sum W_i e^{d_k (l_i - m)}, s2_k = sum W_i
// e^{2 d_k (l_i - m)} for the current candidates (nothing when the search
// is done).
__global__ void __launch_bounds__(kPB) k_cess_pass(const double *l,
    const PDev *pd, const double *ll, long long n, const CessDev *cd,
    double *part)
{
    __shared__ double dk[kCand];
    __shared__ int nk;
    if (threadIdx.x == 0)
        nk = cd->done ? 0 : cd->ncand;
    if (threadIdx.x < kCand)
        dk[threadIdx.x] = cd->cand[threadIdx.x];
    __syncthreads();
    const int nc = nk;
    if (nc == 0)
        return;
    const double m = cd->m;
    const int equal = pd->equal;
    const double c = pd->c;
    const double lw_equal = pd->log_equal;
    double s1[kCand], s2[kCand];
#pragma unroll
    for (int k = 0; k < kCand; ++k) {
        s1[k] = 0.0;
        s2[k] = 0.0;
    }
    const long long base = static_cast<long long>(blockIdx.x) * kPTile;
    for (int j = 0; j < kPItems; ++j) {
        const long long i = base + j * kPB + threadIdx.x;
        if (i >= n)
            continue;
        const double x = ll[i] - m;
        if (x != x)
            continue;   // NaN
        const double W = exp(equal ? lw_equal : l[i] - c);
#pragma unroll
        for (int k = 0; k < kCand; ++k) {
            if (k < nc) {
                // e^{0 ll} = 1 also for ll = -inf
                const double e = dk[k] == 0.0 ? 1.0 : exp(dk[k] * x);
                s1[k] += W * e;
                s2[k] += W * e * e;
            }
        }
    }
    __shared__ double lds[kWaves];
    for (int k = 0; k < nc; ++k) {
        const double a = block_reduce(s1[k], OpAddD(), lds);
        const double b = block_reduce(s2[k], OpAddD(), lds);
        if (threadIdx.x == 0) {
            part[(2 * static_cast<long long>(blockIdx.x)) * kCand + k] = a;
            part[(2 * static_cast<long long>(blockIdx.x) + 1) * kCand + k] = b;
        }
    }
}

// Merges the pass: CESS of each candidate; narrows the bracket to the gap
// between the last candidate at or above the target and the next, and
// spreads the next pass's kCand candidates over (lo, hi].
__global__ void __launch_bounds__(kPB) k_cess_merge(const double *part,
    long long nb, CessDev *cd)
{
    __shared__ double cs[kCand];
    __shared__ int nk;
    if (threadIdx.x == 0)
        nk = cd->done ? 0 : cd->ncand;
    __syncthreads();
    const int nc = nk;
    if (nc == 0)
        return;
    __shared__ double lds[kWaves];
    for (int k = 0; k < nc; ++k) {
        double a = 0.0, b = 0.0;
        for (long long j = threadIdx.x; j < nb; j += kPB) {
            a += part[(2 * j) * kCand + k];
            b += part[(2 * j + 1) * kCand + k];
        }
        a = block_reduce(a, OpAddD(), lds);
        b = block_reduce(b, OpAddD(), lds);
        if (threadIdx.x == 0)
            cs[k] = b > 0.0 ? a * a / b : 0.0;
    }
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    const double target = cd->target;
    for (int k = 0; k < nc; ++k)
        cd->cess[k] = cs[k];
    if (nc == 1 && cd->lo == 0.0 && cd->cand[0] == cd->hi) {
        // first pass of a search: the largest increment itself
        if (cs[0] >= target) {
            cd->lo = cd->hi;
            cd->done = 1;
            return;
        }
    } else {
        int j = -1;
        for (int k = 0; k < nc; ++k)
            if (cs[k] >= target)
                j = k;
        const double lo = j >= 0 ? cd->cand[j] : cd->lo;
        const double hi = j + 1 < nc ? cd->cand[j + 1] : cd->hi;
        cd->lo = lo;
        cd->hi = hi;
    }
    const double lo = cd->lo, hi = cd->hi;
    for (int k = 0; k < kCand; ++k)
        cd->cand[k] = k + 1 < kCand ? lo + (hi - lo) * static_cast<double>(k + 1) / kCand : hi;
    cd->ncand = kCand;
}

static int cess_buffers(ParticleSet *p, long long nb)
{
    if (p->cess_buf)
        return kOk;
    // CessDev, then nb block maxima, then 2 kCand partials per block
    const size_t bytes = sizeof(CessDev) + sizeof(double) * (nb + 2 * kCand * nb);
    VSMCH_HIP(hipMalloc(&p->cess_buf, bytes));
    return kOk;
}

static int cess_run(ParticleSet *p, const double *ll, double d, double target,
    int solve, int rounds, CessDev *h)
{
    if (!ll) {
        set_error("cess: log-likelihoods are null");
        return kErrInvalid;
    }
    const long long n = p->n;
    const long long nb = (n + kPTile - 1) / kPTile;
    int rc = cess_buffers(p, nb);
    if (rc != kOk)
        return rc;
    hipStream_t st = p->ws.stream;
    CessDev *cd = static_cast<CessDev *>(p->cess_buf);
    double *bmax = reinterpret_cast<double *>(cd + 1);
    double *part = bmax + nb;
    hipLaunchKernelGGL(k_ll_max, dim3(nb), dim3(kPB), 0, st, ll, n, bmax);
    hipLaunchKernelGGL(k_cess_start, dim3(1), dim3(kPB), 0, st,
        static_cast<const double *>(bmax), nb, d, target, solve, cd);
    // pass 0 evaluates d alone; each further pass kCand candidates
    for (int r = 0; r <= (solve ? rounds : 0); ++r) {
        hipLaunchKernelGGL(k_cess_pass, dim3(nb), dim3(kPB), 0, st,
            static_cast<const double *>(p->logw), static_cast<const PDev *>(p->pd), ll, n,
            static_cast<const CessDev *>(cd), part);
        hipLaunchKernelGGL(k_cess_merge, dim3(1), dim3(kPB), 0, st,
            static_cast<const double *>(part), nb, cd);
    }
    VSMCH_HIP(hipGetLastError());
    VSMCH_HIP(hipMemcpyAsync(h, cd, sizeof(CessDev), hipMemcpyDeviceToHost, st));
    VSMCH_HIP(hipStreamSynchronize(st));
    return kOk;
}

int particle_cess(ParticleSet *p, const double *ll, double d, double *out)
{
    if (!(d >= 0.0) || !(d < INFINITY)) {
        set_error("cess: the increment must be finite and >= 0");
        return kErrInvalid;
    }
    CessDev h{};
    const int rc = cess_run(p, ll, d, 0.0, 0, 0, &h);
    if (rc == kOk)
        *out = h.cess[0];
    return rc;
}

int particle_cess_solve(ParticleSet *p, const double *ll, double target,
    double dmax, int rounds, double *out)
{
    if (!(target > 0.0 && target <= 1.0) || !(dmax > 0.0 && dmax < INFINITY) ||
        rounds < 1 || rounds > 64) {
        set_error("cess_solve: need 0 < target <= 1, 0 < dmax < inf, 1 <= rounds <= 64");
        return kErrInvalid;
    }
    CessDev h{};
    const int rc = cess_run(p, ll, dmax, target, 1, rounds, &h);
    if (rc == kOk)
        *out = h.lo;
    return rc;
}

} // namespace vsmch
