// particle.hip -- device particle set: weights, fused PF move, monitor.
//
// Reference: Weight (include/vsmc/core/weight.hpp:99-229), Particle::
// resample (include/vsmc/core/particle.hpp:170-197), Monitor::eval
// (include/vsmc/core/monitor.hpp:242-268), Sampler::do_iter / do_resample
// (include/vsmc/core/sampler.hpp:746-778), and the PF model of
// user_guide/cpp/pf_seq.cpp / example/pf/include/pf.hpp.
//
// Log-weights are kept as l_i with an offset c such that the normalised
// weight is W_i = exp(l_i - c); after set_equal the flag `equal` stands for
// W_i = 1/N without touching l. One pass over the particles writes l and
// produces per-block (max, sum exp, sum exp^2) partials; a one-block kernel
// merges them into lmax, the sum, the ESS and the log-evidence increment.
#include "particle.hpp"
#include "mlp.hpp"
#include "../../include/vsmch/device_rng.hpp"

#include <utility>

namespace vsmch {

constexpr int kPB = 256;              // threads per particle block
constexpr int kPItems = 16;           // particles per thread
constexpr int kPTile = kPB * kPItems; // particles per block
// k_prepare writes the resampling fold's tile prefix from these blocks'
// weight partials: one particle block is one fold tile.
static_assert(kPTile == kTile, "particle block must equal the fold tile");

struct Part3 {
    double m, s, q;
};

// Online merge of (max, sum exp(x - max), sum exp(2 (x - max))).
__device__ inline Part3 part_merge(Part3 a, Part3 b)
{
    if (!(a.m > -INFINITY))
        return b;
    if (!(b.m > -INFINITY))
        return a;
    if (a.m >= b.m) {
        const double r = exp(b.m - a.m);
        return Part3{a.m, a.s + b.s * r, a.q + b.q * r * r};
    }
    const double r = exp(a.m - b.m);
    return Part3{b.m, b.s + a.s * r, b.q + a.q * r * r};
}

struct OpPart {
    __device__ Part3 operator()(Part3 a, Part3 b) const { return part_merge(a, b); }
};

__device__ inline Part3 part_of(double l)
{
    return Part3{l, (l > -INFINITY) ? 1.0 : 0.0, (l > -INFINITY) ? 1.0 : 0.0};
}

__device__ inline void block_partials_at(Part3 p, Part3 *dst)
{
    __shared__ Part3 lds[kWaves];
    p = block_reduce(p, OpPart(), lds);
    if (threadIdx.x == 0)
        *dst = p;
}

__device__ inline void block_partials(Part3 p, Part3 *out)
{
    block_partials_at(p, out + blockIdx.x);
}

// Partials (max, sum exp, sum exp^2) of k log-weights held by one thread:
// the maximum first, then one exp per value. -inf and NaN contribute
// nothing, as in part_of / part_merge.
template <int K>
__device__ inline Part3 part_of_array(const double (&v)[K])
{
    double mx = -INFINITY;
    for (int k = 0; k < K; ++k)
        mx = v[k] > mx ? v[k] : mx;
    Part3 p{mx, 0.0, 0.0};
    if (mx > -INFINITY) {
        for (int k = 0; k < K; ++k) {
            if (v[k] > -INFINITY) {
                const double r = exp(v[k] - mx);
                p.s += r;
                p.q += r * r;
            }
        }
    }
    return p;
}

// A log-weight that came out NaN (a non-finite state: a float uniform of
// exactly 0 gives an infinite Box-Muller radius, probability 2^-24 per
// draw) becomes -inf: the particle carries no weight and the next
// resampling replaces it. Zero-weight particles also stay out of the
// weighted sums (0 * inf would be NaN).
__device__ inline double dead_if_nan(double l)
{
    return l == l ? l : -INFINITY;
}

// log W_i of the current weights.
__device__ inline double logw_of(const double *l, long long i, const PDev *pd)
{
    return pd->equal ? pd->log_equal : l[i] - pd->c;
}

// Weight updates from src[i * stride] (weight.hpp:105-171): mode 0 set_log
// (l = src), 1 add_log (l = log W + src), 3 set (l = log src), 4 mul
// (l = log W + log src). Linear weights are kept as their logarithms.
__global__ void __launch_bounds__(kPB) k_weight_pass(double *l,
    const double *src, long long n, int mode, long long stride, const PDev *pd,
    Part3 *part)
{
    const long long base = static_cast<long long>(blockIdx.x) * kPTile;
    const bool old = mode == 1 || mode == 4;   // the update reads log W
    const int equal = pd->equal;
    const double log_equal = pd->log_equal;
    const double c = pd->c;
    // clamped, unconditional loads, all in flight before their first use
    double a[kPItems], lo[kPItems];
#pragma unroll
    for (int j = 0; j < kPItems; ++j) {
        const long long i = base + j * kPB + threadIdx.x;
        const long long ic = i < n ? i : n - 1;
        a[j] = src[ic * stride];
        lo[j] = old && !equal ? l[ic] : 0.0;
    }
    double v[kPItems];
#pragma unroll
    for (int j = 0; j < kPItems; ++j) {
        keep_loaded(a[j]);
        keep_loaded(lo[j]);
        const long long i = base + j * kPB + threadIdx.x;
        const double in = mode >= 3 ? log(a[j]) : a[j];
        const double lw = equal ? log_equal : lo[j] - c;
        v[j] = old ? lw + in : in;
        if (i < n)
            l[i] = v[j];
        else
            v[j] = -INFINITY;
    }
    block_partials(part_of_array(v), part);
}

// Merges block partials; mode 0: set_log, 1: add_log, 2: set_equal.
// Also prepares the resampling parameters of this step (weights from l)
// and records the history entry `t` when hist is non-null. kThreads: the
// block size it runs with.
// guard: the step was a learned move; a collapsed ESS sets pd->fallback.
template <int kThreads = kPB>
__device__ inline void weight_merge(const Part3 *part, long long nblocks,
    long long n, int mode, PDev *pd, int guard = 0);

// One block of kMergeThreads: four times the threads of a particle block,
// so each thread merges a quarter of the partials (the merge is a chain of
// dependent loads and exponentials per thread, latency-bound).
constexpr int kMergeThreads = 1024;

__global__ void __launch_bounds__(kMergeThreads) k_weight_finalize(const Part3 *part,
    long long nblocks, long long n, int mode, PDev *pd, int guard = 0)
{
    weight_merge<kMergeThreads>(part, nblocks, n, mode, pd, guard);
}

// Block reduction over kThreads threads, the result broadcast (lds: one
// entry per wave).
template <int kThreads, typename T, typename Op>
__device__ inline T block_reduce_n(T v, Op op, T *lds)
{
    constexpr int nw = kThreads / kWave;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    v = wave_reduce(v, op);
    if (lane == 0)
        lds[wave] = v;
    __syncthreads();
    T r = lds[0];
#pragma unroll
    for (int w = 1; w < nw; ++w)
        r = op(r, lds[w]);
    __syncthreads();
    return r;
}

// Sums of K values per thread over kThreads threads in one reduction (one
// pair of barriers for all K), broadcast; lds: K entries per wave.
template <int kThreads, int K>
__device__ inline void block_sum_n(double (&v)[K], double *lds)
{
    constexpr int nw = kThreads / kWave;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k)
        v[k] = wave_reduce(v[k], OpAddD());
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            lds[wave * K + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double r = lds[k];
#pragma unroll
        for (int w = 1; w < nw; ++w)
            r += lds[w * K + k];
        v[k] = r;
    }
    __syncthreads();
}

// Partials a merge thread holds in registers: thread t has partials
// t kMergePer .. t kMergePer + kMergePer - 1, loaded together.
constexpr int kMergePer = 4;

// (max, S, Q) of the block partials part[0, nblocks) over kThreads threads,
// broadcast to every thread: the maximum first, then independent scaled sums
// (no serial chain of exponentials through a running maximum). With nblocks
// <= kThreads kMergePer every thread loads its partials at once (clamped,
// unconditional loads) and keeps them: held is set and sb[k] = s_b e^{m_b -
// max} of its k-th partial (0 past nblocks). Otherwise strided loops.
// extra (when non-null): two more per-thread sums, reduced in the same
// pass (in place). dlds: four entries per wave.
template <int kThreads>
__device__ inline Part3 merge_partials(const Part3 *part, long long nblocks,
    double (&sb)[kMergePer], bool &held, double *dlds, double *extra = nullptr)
{
    if (nblocks <= static_cast<long long>(kThreads) * kMergePer) {
        held = true;
        const long long b0 = static_cast<long long>(threadIdx.x) * kMergePer;
        double m[kMergePer], s[kMergePer], q[kMergePer];
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            const long long b = b0 + k < nblocks ? b0 + k : nblocks - 1;
            m[k] = part[b].m;
            s[k] = part[b].s;
            q[k] = part[b].q;
        }
        double mx = -INFINITY;
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            keep_loaded(m[k]);
            keep_loaded(s[k]);
            keep_loaded(q[k]);
            if (b0 + k < nblocks)
                mx = m[k] > mx ? m[k] : mx;
        }
        mx = block_reduce_n<kThreads>(mx, OpMaxD(), dlds);
        double S = 0.0, Q = 0.0;
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            double r = 0.0;
            if (b0 + k < nblocks && m[k] > -INFINITY && mx > -INFINITY)
                r = exp(m[k] - mx);
            sb[k] = s[k] * r;
            S += sb[k];
            Q += q[k] * r * r;
        }
        double v[4] = {S, Q, extra ? extra[0] : 0.0, extra ? extra[1] : 0.0};
        block_sum_n<kThreads, 4>(v, dlds);
        if (extra) {
            extra[0] = v[2];
            extra[1] = v[3];
        }
        return Part3{mx, v[0], v[1]};
    }
    held = false;
    double mx = -INFINITY;
    for (long long b = threadIdx.x; b < nblocks; b += kThreads) {
        const double m = part[b].m;
        mx = m > mx ? m : mx;
    }
    mx = block_reduce_n<kThreads>(mx, OpMaxD(), dlds);
    double S = 0.0, Q = 0.0;
    if (mx > -INFINITY) {
        for (long long b = threadIdx.x; b < nblocks; b += kThreads) {
            const Part3 p = part[b];
            if (p.m > -INFINITY) {
                const double r = exp(p.m - mx);
                S += p.s * r;
                Q += p.q * r * r;
            }
        }
    }
    double v[4] = {S, Q, extra ? extra[0] : 0.0, extra ? extra[1] : 0.0};
    block_sum_n<kThreads, 4>(v, dlds);
    if (extra) {
        extra[0] = v[2];
        extra[1] = v[3];
    }
    return Part3{mx, v[0], v[1]};
}

// The weight scalars from a merged (max, S, Q): thread 0 only.
__device__ inline void weight_merge_store(Part3 acc, long long n, int mode,
    PDev *pd, int guard);

// The body of k_weight_finalize (one block of kThreads).
template <int kThreads>
__device__ inline void weight_merge(const Part3 *part, long long nblocks,
    long long n, int mode, PDev *pd, int guard)
{
    __shared__ double dlds[4 * (kThreads / kWave)];
    Part3 acc{-INFINITY, 0.0, 0.0};
    if (mode != 2) {
        double sb[kMergePer];
        bool held;
        acc = merge_partials<kThreads>(part, nblocks, sb, held, dlds);
    }
    if (threadIdx.x == 0)
        weight_merge_store(acc, n, mode, pd, guard);
}

__device__ inline void weight_merge_store(Part3 acc, long long n, int mode,
    PDev *pd, int guard)
{
    if (mode == 2) {
        pd->equal = 1;
        pd->log_equal = -log(static_cast<double>(n));
        pd->ess = static_cast<double>(n);
        pd->lmax = pd->log_equal;
        pd->sum = 1.0;
        pd->c = pd->log_equal;
        return;
    }
    pd->equal = 0;
    pd->lmax = acc.m;
    pd->sum = acc.s;
    pd->c = acc.m + log(acc.s);
    // ESS = 1 / sum W^2 = S^2 / Q (weight.hpp:43-47)
    pd->ess = acc.s * acc.s / acc.q;
    // evidence increment: log sum_i W_prev,i exp(inc_i) = c for add_log
    // (l = log W_prev + inc); set_log assumes equal previous weights.
    const double inc = mode == 0 ? pd->c - log(static_cast<double>(n)) : pd->c;
    pd->logz_inc = inc;
    pd->logz += inc;
    if (mode == 0)
        pd->fallback = 0;
    if (guard && !(pd->ess >= kLearnedCollapse * static_cast<double>(n)))
        pd->fallback = 1;
}

// Resampling decision and event parameters (Particle::resample,
// particle.hpp:172-187; Sampler::do_resample records ESS first,
// sampler.hpp:772-778).
__global__ void k_decide(PDev *pd, ResampleDev *prm, long long n,
    double threshold, long long t, double *ess_hist, int *rs_hist)
{
    const double ess = pd->ess;
    const int active = ess < threshold * static_cast<double>(n) ? 1 : 0;
    if (ess_hist)
        ess_hist[t] = ess;
    if (rs_hist)
        rs_hist[t] = active;
    pd->resampled = active;
    prm->active = active;
    prm->M = n;
    prm->nsample = n;
    prm->ndraw = n;
    prm->key = pd->rkey;
    prm->word0 = pd->rword;
    prm->err = 0;
    prm->integ_sum = 0;
    prm->words_used = 0;
    prm->equal = pd->equal;
    if (pd->equal) {
        prm->lmax = pd->log_equal;
        prm->inv_sum = 1.0;
    } else {
        prm->lmax = pd->lmax;
        prm->inv_sum = 1 / pd->sum;
    }
}

// Decision and event parameters as k_decide, plus what the resampling
// passes would otherwise recompute from the weights: delta = fl(1/N), the
// systematic start u0 = U delta (u01_sequence.hpp:182-187, :231-236), and
// the approximate exclusive prefix of the normalised weights over tiles of
// kPTile particles, from the weight pass's block partials
// (sum_b exp(m_b - lmax) / sum, one term per block).
// The decision and draw parameters of a systematic / stratified event
// (one thread): returns whether the step resamples.
__device__ inline int decide_event(PDev *pd, ResampleDev *prm, long long n,
    double threshold, int scheme)
{
    const double ess = pd->ess;
    const int active = ess < threshold * static_cast<double>(n) ? 1 : 0;
    pd->resampled = active;
    prm->active = active;
    prm->M = n;
    prm->nsample = n;
    prm->ndraw = n;
    prm->key = pd->rkey;
    prm->word0 = pd->rword;
    prm->err = 0;
    prm->integ_sum = 0;
    prm->words_used = 0;
    prm->equal = pd->equal;
    prm->lmax = pd->equal ? pd->log_equal : pd->lmax;
    prm->inv_sum = pd->equal ? 1.0 : 1 / pd->sum;
    const double delta = 1 / static_cast<double>(n);
    prm->delta = delta;
    prm->u0 = (scheme == kSystematic) ?
        u01_u32d(philox_word(pd->rkey, pd->rword)) * delta : 0.0;
    return active;
}

__global__ void __launch_bounds__(kMergeThreads) k_prepare(PDev *pd, ResampleDev *prm,
    const Part3 *part, long long nb, long long n, double threshold,
    long long t, double *ess_hist, int *rs_hist, int scheme, double *tpre,
    int *fpred)
{
    __shared__ int act;
    __shared__ double lds[kMergeThreads / kWave];
    if (threadIdx.x == 0) {
        if (ess_hist)
            ess_hist[t] = pd->ess;
        const int active = decide_event(pd, prm, n, threshold, scheme);
        if (rs_hist)
            rs_hist[t] = active;
        act = active;
    }
    __syncthreads();
    if (!act)
        return;
    const int equal = pd->equal;
    const double lmax = pd->lmax;
    const double inv = equal ? 1.0 : 1 / pd->sum;
    const double inv_n = 1 / static_cast<double>(n);
    block_scan_chunked_n<kMergeThreads, 4>(nb,
        [&](long long b) {
            if (equal) {
                const long long rem = n - b * kPTile;
                return static_cast<double>(rem < kPTile ? rem : kPTile) * inv_n;
            }
            const Part3 q = part[b];
            return q.s > 0.0 ? q.s * exp(q.m - lmax) * inv : 0.0;
        },
        [&](long long b, double ex) { tpre[b] = ex; }, OpAddD(), 0.0, lds);
    __shared__ int npred;
    list_predicted(tpre, nb, fpred, &npred);
}

// A filter step's bookkeeping after its move, one block of kMergeThreads:
// what k_weight_finalize, k_moments_finalize, the log-evidence record and
// k_prepare (k_decide) do in turn, in one launch. Merges the move's weight
// partials (mode 0: set_log at the initialisation, 1: add_log), records the
// log-evidence, the previous step's monitor from the move's partials
// (pos_prev, when given), the ESS before resampling and the decision
// (sampler.hpp:772-778), and for systematic / stratified events (prep) the
// draw parameters and the approximate tile prefix. The move has read the
// ancestors a previous event left, so anc_valid is cleared (and with swap
// the current buffer pd->cur changes when it was set).
__global__ void __launch_bounds__(kMergeThreads) k_step_finalize(
    const Part3 *part, long long nb, long long n, int mode, PDev *pd,
    ResampleDev *prm, double threshold, long long h, double *ess_hist,
    int *rs_hist, double *logz_hist, const double *mpart, double *pos_prev,
    int scheme, int prep, double *tpre, int *fpred, int swap)
{
    __shared__ int act;
    __shared__ int sh_equal;
    __shared__ double sh_lmax, sh_inv;
    __shared__ double lds[4 * (kMergeThreads / kWave)];
    // the "pos" monitor's partials load with the weight partials (a thread's
    // four consecutive blocks, clamped, unconditional loads)
    const bool mheld = pos_prev && nb <= static_cast<long long>(kMergeThreads) * kMergePer;
    double mp[2][kMergePer];
    if (mheld) {
        const long long b0 = static_cast<long long>(threadIdx.x) * kMergePer;
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            const long long b = b0 + k < nb ? b0 + k : nb - 1;
            mp[0][k] = mpart[b * 8 + 0];
            mp[1][k] = mpart[b * 8 + 1];
        }
    }
    // the particle set's scalars the step reads, loaded before the merge so
    // that thread 0's tail below runs from registers
    double logz0 = 0.0;
    unsigned int rkey = 0;
    unsigned long long rword = 0;
    if (threadIdx.x == 0) {
        logz0 = pd->logz;
        rkey = pd->rkey;
        rword = pd->rword;
    }
    // the monitor's per-thread sums, reduced with the weight sums
    double ms[2] = {0.0, 0.0};
    if (pos_prev) {
        for (int d = 0; d < 2; ++d) {
            if (mheld) {
                const long long b0 = static_cast<long long>(threadIdx.x) * kMergePer;
#pragma unroll
                for (int k = 0; k < kMergePer; ++k) {
                    keep_loaded(mp[d][k]);
                    ms[d] += b0 + k < nb ? mp[d][k] : 0.0;
                }
            } else {
                for (long long b = threadIdx.x; b < nb; b += kMergeThreads)
                    ms[d] += mpart[b * 8 + d];
            }
        }
    }
    Part3 acc{-INFINITY, 0.0, 0.0};
    double sb[kMergePer];
    bool held = false;
    if (mode != 2) {
        acc = merge_partials<kMergeThreads>(part, nb, sb, held, lds, ms);
    } else if (pos_prev) {
        block_sum_n<kMergeThreads, 2>(ms, lds);
    }
    if (pos_prev && threadIdx.x == 0) {
        pos_prev[0] = ms[0];
        pos_prev[1] = ms[1];
    }
    if (threadIdx.x == 0) {
        // weight_merge_store and decide_event on registers: the same
        // arithmetic, every result stored once
        const double nd = static_cast<double>(n);
        const int equal = mode == 2 ? 1 : 0;
        const double log_equal = -log(nd);
        const double lmax = equal ? log_equal : acc.m;
        const double sum = equal ? 1.0 : acc.s;
        const double c = equal ? log_equal : acc.m + log(acc.s);
        const double ess = equal ? nd : acc.s * acc.s / acc.q;
        const double inc = mode == 0 ? c - log(nd) : c;
        const double logz = equal ? logz0 : logz0 + inc;
        pd->equal = equal;
        pd->lmax = lmax;
        pd->sum = sum;
        pd->c = c;
        pd->ess = ess;
        if (equal) {
            pd->log_equal = log_equal;
        } else {
            pd->logz_inc = inc;
            pd->logz = logz;
        }
        if (mode == 0)
            pd->fallback = 0;
        // swap: the move could read through the ancestors (cv_iterate's two
        // buffers); when it did, it wrote the other buffer
        if (swap && pd->anc_valid)
            pd->cur ^= 1;
        pd->anc_valid = 0;
        if (logz_hist)
            logz_hist[h] = logz;
        if (ess_hist)
            ess_hist[h] = ess;
        const int active = ess < threshold * nd ? 1 : 0;
        pd->resampled = active;
        prm->active = active;
        prm->M = n;
        prm->nsample = n;
        prm->ndraw = n;
        prm->key = rkey;
        prm->word0 = rword;
        prm->err = 0;
        prm->integ_sum = 0;
        prm->words_used = 0;
        prm->equal = equal;
        prm->lmax = lmax;
        prm->inv_sum = equal ? 1.0 : 1 / sum;
        const double delta = 1 / nd;
        prm->delta = delta;
        prm->u0 = (scheme == kSystematic) ? u01_u32d(philox_word(rkey, rword)) * delta : 0.0;
        if (rs_hist)
            rs_hist[h] = active;
        act = active;
        sh_equal = equal;
        sh_lmax = lmax;
        sh_inv = equal ? 1.0 : 1 / sum;
    }
    __syncthreads();
    if (!act || !prep)
        return;
    const int equal = sh_equal;
    const double lmax = sh_lmax;
    const double inv = sh_inv;
    const double inv_n = 1 / static_cast<double>(n);
    __shared__ int npred;
    if (held && !equal) {
        // the approximate tile prefix from the scaled sums already held, and
        // the predicted crossing tiles (fold::predicted_crossing, on the
        // same values: a tile's start as written, the next thread's first
        // through LDS)
        __shared__ double first[kMergeThreads];
        const long long b0 = static_cast<long long>(threadIdx.x) * kMergePer;
        double v[kMergePer];
        double loc = 0.0;
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            v[k] = sb[k] * inv;
            loc += v[k];
        }
        if (threadIdx.x == 0)
            npred = 0;
        double tot;
        double run = block_exclusive_scan_n<kMergeThreads>(loc, OpAddD(), 0.0, lds, &tot);
        double pre[kMergePer + 1];
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            pre[k] = run;
            if (b0 + k < nb)
                tpre[b0 + k] = run;
            run += v[k];
        }
        first[threadIdx.x] = pre[0];
        __syncthreads();
        pre[kMergePer] = threadIdx.x + 1 < kMergeThreads ? first[threadIdx.x + 1] : run;
#pragma unroll
        for (int k = 0; k < kMergePer; ++k) {
            const long long b = b0 + k;
            if (b < nb && (b == 0 || b + 1 >= nb ||
                    fold::sum_tag(pre[k]) != fold::sum_tag(pre[k + 1]))) {
                const int q = atomicAdd(&npred, 1);
                if (q < kFePred)
                    fpred[1 + q] = static_cast<int>(b);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            fpred[0] = npred;
        return;
    } else {
        block_scan_chunked_n<kMergeThreads, 4>(nb,
            [&](long long b) {
                if (equal) {
                    const long long rem = n - b * kPTile;
                    return static_cast<double>(rem < kPTile ? rem : kPTile) * inv_n;
                }
                const Part3 q = part[b];
                return q.s > 0.0 ? q.s * exp(q.m - lmax) * inv : 0.0;
            },
            [&](long long b, double ex) { tpre[b] = ex; }, OpAddD(), 0.0, lds);
    }
    list_predicted(tpre, nb, fpred, &npred);
}

// cv_iterate starts from the particle set's own buffer (a kernel, not a
// memset, so that a captured sequence replays it).
__global__ void k_own_buffer(PDev *pd)
{
    pd->cur = 0;
}

// The state after the last move of cv_iterate, into the particle set's own
// buffer home: particle i from the current buffer (pd->cur: home or alt) at
// anc[i] when the last event left its ancestors (pd->anc_valid), else at i.
// With the current buffer home it runs in place and writes only the slots
// whose ancestor is another particle: such an ancestor has copies, so it
// keeps its own slot and is never written.
template <typename Real>
__global__ void __launch_bounds__(kPB) k_settle(const Real *alt, Real *home,
    const int *anc, const PDev *pd, long long n, int dim)
{
    const long long i = static_cast<long long>(blockIdx.x) * kPB + threadIdx.x;
    if (i >= n)
        return;
    const Real *const cur = pd->cur ? alt : home;
    const long long a = pd->anc_valid ? static_cast<long long>(anc[i]) : i;
    if (cur == home && a == i)
        return;
    for (int d = 0; d < dim; ++d)
        home[d * n + i] = cur[d * n + a];
}

// Weighted moments r_d = sum_i W_i x_d(i) of components [c0, c0 + dim)
// (dim <= 8); per-block partials of sum exp(l - lmax) x and the merge.
template <typename Real>
__global__ void __launch_bounds__(kPB) k_moments(const Real *state,
    const double *l, long long n, int c0, int dim, const PDev *pd,
    double *part)
{
    const long long base = static_cast<long long>(blockIdx.x) * kPTile;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < kPItems; ++j) {
        const long long i = base + j * kPB + threadIdx.x;
        if (i < n) {
            const double w = exp(logw_of(l, i, pd));
            if (w > 0.0)
                for (int d = 0; d < dim; ++d)
                    acc[d] += w * static_cast<double>(state[(c0 + d) * n + i]);
        }
    }
    __shared__ double lds[kWaves];
    for (int d = 0; d < dim; ++d) {
        const double v = block_reduce(acc[d], OpAddD(), lds);
        if (threadIdx.x == 0)
            part[blockIdx.x * 8 + d] = v;
    }
}

__global__ void __launch_bounds__(kPB) k_moments_finalize(const double *part,
    long long nblocks, int dim, double *out)
{
    __shared__ double lds[kWaves];
    for (int d = 0; d < dim; ++d) {
        double acc = 0.0;
        for (long long b = threadIdx.x; b < nblocks; b += kPB)
            acc += part[b * 8 + d];
        acc = block_reduce(acc, OpAddD(), lds);
        if (threadIdx.x == 0)
            out[d] = acc;
    }
}

// Normalised linear weights, the numbers the resampling passes use:
// W_i = e^(l_i - lmax) * (1 / sum) (e^x by wexp), the reference's order of
// operations (weight.hpp:200-222: subtract the maximum, exponentiate, scale
// by the reciprocal of the sum).
__global__ void k_weight_read(const double *l, long long n, const PDev *pd,
    double *W)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        W[i] = pd->equal ? 1.0 / static_cast<double>(n) :
                           (1 / pd->sum) * wexp(l[i] - pd->lmax);
}

// ------------------------------------------------------------ CV model

template <typename Real>
__device__ inline Real radd(Real a, Real b);
template <>
__device__ inline double radd<double>(double a, double b) { return rn_add(a, b); }
template <>
__device__ inline float radd<float>(float a, float b) { return rn_add(a, b); }
template <typename Real>
__device__ inline Real rmul(Real a, Real b);
template <>
__device__ inline double rmul<double>(double a, double b) { return rn_mul(a, b); }
template <>
__device__ inline float rmul<float>(float a, float b) { return rn_mul(a, b); }

// log-likelihood of example/pf/include/pf.hpp:78-87 (Student-t with the
// reference's dropped constants), optionally with the constants, or
// Gaussian. Evaluated in Real.
//
// FP64 state follows the reference's operations one for one. FP32 state is
// a GPU-only configuration with no reference counterpart to match bit for
// bit; there the two logarithms use the hardware log2 (v_log_f32, about
// 1e-7 absolute error on log-weights whose scale is 1) and the divisions
// become products with precomputed reciprocals.
__device__ inline double ll_log(double v) { return log(v); }
__device__ inline float ll_log(float v) { return __logf(v); }
__device__ inline double ll_div(double v, double d, double) { return v / d; }
__device__ inline float ll_div(float v, double, double inv)
{
    return v * static_cast<float>(inv);
}

template <typename Real>
__device__ inline Real cv_loglik(const CvModelDev &m, Real x, Real y, Real ox,
    Real oy)
{
    if (m.obs_model == 2) {
        const Real rx = ll_div(static_cast<Real>(x - ox), m.obs_scale, m.inv_scale);
        const Real ry = ll_div(static_cast<Real>(y - oy), m.obs_scale, m.inv_scale);
        return static_cast<Real>(-0.5) * (rx * rx + ry * ry) +
            static_cast<Real>(m.ll_const);
    }
    const Real scale = static_cast<Real>(m.inv_scale);
    const Real nu = static_cast<Real>(m.obs_df);
    Real lx = rmul(scale, static_cast<Real>(x - ox));
    Real ly = rmul(scale, static_cast<Real>(y - oy));
    lx = ll_log(radd(static_cast<Real>(1), ll_div(rmul(lx, lx), m.obs_df, m.inv_df)));
    ly = ll_log(radd(static_cast<Real>(1), ll_div(rmul(ly, ly), m.obs_df, m.inv_df)));
    const Real c = rmul(static_cast<Real>(-0.5), radd(nu, static_cast<Real>(1)));
    Real ll = rmul(c, radd(lx, ly));
    if (m.obs_model == 1)
        ll = ll + static_cast<Real>(m.ll_const);
    return ll;
}

// PFInit::eval_sp (user_guide/cpp/pf_seq.cpp:70-81): 4 normals from the
// first block of the particle's stream: x, y from one Box-Muller pair with
// sd_pos0, vx, vy from the next with sd_vel0; then the log-likelihood.
// PFMove::eval_sp (:111-122) at step t uses block t + 1:
// x += sd_pos z0 + dt vx, y += sd_pos z1 + dt vy, vx += sd_vel z2,
// vy += sd_vel z3, then log-likelihood; weights by add_log.
// Also accumulates the monitor moments of the state *before* the move
// (the previous step's post-resampling "pos" record) when mon != null.
template <typename Real>
struct alignas(4 * sizeof(Real)) Vec4 {
    Real v[4];
};

// Four log-weights stored nontemporally: the move's weights are next read
// by the resampling passes only, while its state, stored normally, stays in
// the last-level cache for the next move's gather through the ancestors
// (measured at N = 2^24: both kinds nontemporal speed the plain move by
// 9 %, but slow the move after an event by 18 %). Two 16-byte halves, each
// made opaque first: the optimiser would otherwise split a store of a
// vector it sees assembled from scalars into scalar stores and drop the
// nontemporal flag.
// Four log-weights loaded nontemporally (read once a step, stored
// nontemporally by the previous move).
__device__ inline void load_nt(const double *p, double (&a)[4])
{
    typedef double R2 __attribute__((ext_vector_type(2)));
    const R2 v0 = __builtin_nontemporal_load(reinterpret_cast<const R2 *>(p));
    const R2 v1 = __builtin_nontemporal_load(reinterpret_cast<const R2 *>(p) + 1);
    a[0] = v0.x;
    a[1] = v0.y;
    a[2] = v1.x;
    a[3] = v1.y;
}

__device__ inline void store_nt(double *p, const double (&a)[4])
{
    typedef double R2 __attribute__((ext_vector_type(2)));
    R2 v0 = {a[0], a[1]}, v1 = {a[2], a[3]};
    asm volatile("" : "+v"(v0));
    asm volatile("" : "+v"(v1));
    __builtin_nontemporal_store(v0, reinterpret_cast<R2 *>(p));
    __builtin_nontemporal_store(v1, reinterpret_cast<R2 *>(p) + 1);
}

// seed_key(s, i + 1) from k = seed_key(s, i): keys run 1 .. 2^32 - 1.
__device__ inline uint32_t next_key(uint32_t k)
{
    return k == 0xFFFFFFFFu ? 1u : k + 1u;
}

// A Box-Muller pair of the model's noise. FP64: the reference's
// NormalDistribution formula. FP32: the same radius; the angle's sine and
// cosine come from sincospif(2u), i.e. sin(2 pi u) without first rounding
// 2 pi u to float (no reference counterpart: FP32 state is GPU-only).
template <typename Real>
__device__ inline normal_pair<Real> cv_normals(uint32_t wr, uint32_t wa)
{
    return box_muller<Real>(wr, wa);
}
// FP32 state: the hardware's single-precision log2, square root and sine /
// cosine (v_log_f32, v_sqrt_f32, v_sin_f32, v_cos_f32; the angle argument is
// in revolutions, so sin(2 pi a) is one instruction). Each is within a few
// FP32 ulps; the library routines cost tens of instructions per call, and
// the move is instruction-bound.
template <>
__device__ inline normal_pair<float> cv_normals<float>(uint32_t wr, uint32_t wa)
{
    constexpr float kTwoLn2 = 1.38629436f;   // 2 ln 2: -2 ln u = -2 ln2 log2 u
    const float r = __builtin_amdgcn_sqrtf(-kTwoLn2 * __builtin_amdgcn_logf(u01_u32f(wr)));
    const float a = u01_u32f(wa);
    return normal_pair<float>{r * __builtin_amdgcn_cosf(a), r * __builtin_amdgcn_sinf(a)};
}

// One particle's move (or initialisation) from its Philox block `blk`.
template <typename Real, bool kInit>
__device__ inline void cv_advance(const CvModelDev &m, uint32_t key,
    unsigned long long blk, Real &x, Real &y, Real &vx, Real &vy)
{
    const u32x4 wd = philox_block(key, blk);
    const normal_pair<Real> p0 = cv_normals<Real>(wd.v[0], wd.v[1]);
    const normal_pair<Real> p1 = cv_normals<Real>(wd.v[2], wd.v[3]);
    if (kInit) {
        x = rmul(static_cast<Real>(m.sd_pos0), p0.z0);
        y = rmul(static_cast<Real>(m.sd_pos0), p0.z1);
        vx = rmul(static_cast<Real>(m.sd_vel0), p1.z0);
        vy = rmul(static_cast<Real>(m.sd_vel0), p1.z1);
    } else {
        const Real dt = static_cast<Real>(m.dt);
        x = radd(x, radd(rmul(static_cast<Real>(m.sd_pos), p0.z0), rmul(dt, vx)));
        y = radd(y, radd(rmul(static_cast<Real>(m.sd_pos), p0.z1), rmul(dt, vy)));
        vx = radd(vx, rmul(static_cast<Real>(m.sd_vel), p1.z0));
        vy = radd(vy, rmul(static_cast<Real>(m.sd_vel), p1.z1));
    }
}

// Tile `tile` (kPTile particles) of a filter of n particles whose moved
// state components go to X, Y, VX, VY; writes the tile's weight partial to
// *part_out and, with mon_out, the monitor sums of the pre-move state.
// The pre-move state is read from SX, SY, SVX, SVY (which may be X, Y, VX,
// VY, for a move in place) at particle i, or at its ancestor anc[i] when the
// last event left its ancestors there (pd->anc_valid): such a move also
// performs the event's gather.
template <typename Real, bool kInit, int kB>
__device__ inline void cv_tile(const Real *SX, const Real *SY, const Real *SVX,
    const Real *SVY, const int *anc, Real *X, Real *Y, Real *VX, Real *VY,
    double *__restrict__ l,
    long long n, long long tile, uint32_t seed, long long t, Real ox, Real oy,
    const CvModelDev &m, const PDev *pd, Part3 *part_out, double *mon_out)
{
    const bool mon = mon_out != nullptr;
    const int *A = (anc != nullptr && pd->anc_valid) ? anc : nullptr;
    const long long base = tile * kPTile;
    Part3 acc{-INFINITY, 0.0, 0.0};
    double mx = 0.0, my = 0.0;
    const double lw_equal = pd->log_equal;
    const int equal = pd->equal;
    const double c = pd->c;
    // equal weights (the step after an event): one monitor weight for all
    const double w_equal = mon && equal ? exp(lw_equal) : 0.0;
    const unsigned long long blk = static_cast<unsigned long long>(t) + 1;
    if ((n & 3) == 0 && base + kPTile <= n) {
        // Full tile, every component row 16-byte aligned: each thread takes
        // kR runs of 4 consecutive particles, kB runs at a time; a batch's
        // loads are all issued before its arithmetic so that they overlap.
        constexpr int kR = kPTile / (4 * kPB);
        static_assert(kR % kB == 0, "batch must divide the runs");
        double lv[kR * 4];
#pragma unroll
        for (int r0 = 0; r0 < kR; r0 += kB) {
            Vec4<Real> xs[kB], ys[kB], vxs[kB], vys[kB];
#pragma unroll
            for (int r = 0; r < kB; ++r) {
                const long long i0 = base + ((r0 + r) * kPB + threadIdx.x) * 4;
                if (!kInit) {
                    if (A) {
                        const Vec4<int> a4 = *reinterpret_cast<const Vec4<int> *>(A + i0);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const unsigned a = static_cast<unsigned>(a4.v[q]);
                            xs[r].v[q] = SX[a];
                            ys[r].v[q] = SY[a];
                            vxs[r].v[q] = SVX[a];
                            vys[r].v[q] = SVY[a];
                        }
                    } else {
                        xs[r] = *reinterpret_cast<const Vec4<Real> *>(SX + i0);
                        ys[r] = *reinterpret_cast<const Vec4<Real> *>(SY + i0);
                        vxs[r] = *reinterpret_cast<const Vec4<Real> *>(SVX + i0);
                        vys[r] = *reinterpret_cast<const Vec4<Real> *>(SVY + i0);
                    }
                    if (!equal) {
                        double lq[4];
                        load_nt(l + i0, lq);
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            lv[(r0 + r) * 4 + q] = lq[q] - c;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            lv[(r0 + r) * 4 + q] = lw_equal;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kB; ++r) {
                const long long i0 = base + ((r0 + r) * kPB + threadIdx.x) * 4;
                uint32_t key = seed_key(seed, static_cast<unsigned long long>(i0));
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    double &lq = lv[(r0 + r) * 4 + q];
                    Real x{}, y{}, vx{}, vy{};
                    if (!kInit) {
                        x = xs[r].v[q];
                        y = ys[r].v[q];
                        vx = vxs[r].v[q];
                        vy = vys[r].v[q];
                        if (mon) {
                            const double w = equal ? w_equal : exp(lq);
                            if (w > 0.0) {
                                mx += w * static_cast<double>(x);
                                my += w * static_cast<double>(y);
                            }
                        }
                    }
                    cv_advance<Real, kInit>(m, key, blk, x, y, vx, vy);
                    key = next_key(key);
                    xs[r].v[q] = x;
                    ys[r].v[q] = y;
                    vxs[r].v[q] = vx;
                    vys[r].v[q] = vy;
                    const double ll = static_cast<double>(cv_loglik<Real>(m, x, y, ox, oy));
                    lq = dead_if_nan(kInit ? ll : lq + ll);
                }
                *reinterpret_cast<Vec4<Real> *>(X + i0) = xs[r];
                *reinterpret_cast<Vec4<Real> *>(Y + i0) = ys[r];
                *reinterpret_cast<Vec4<Real> *>(VX + i0) = vxs[r];
                *reinterpret_cast<Vec4<Real> *>(VY + i0) = vys[r];
                double lo[4];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    lo[q] = lv[(r0 + r) * 4 + q];
                store_nt(l + i0, lo);
            }
        }
        acc = part_of_array(lv);
    } else {
        for (int j = 0; j < kPItems; ++j) {
            const long long i = base + j * kPB + threadIdx.x;
            if (i >= n)
                continue;
            const uint32_t key = seed_key(seed, static_cast<unsigned long long>(i));
            Real x{}, y{}, vx{}, vy{};
            double lold = 0.0;
            if (!kInit) {
                const long long a = A ? static_cast<long long>(A[i]) : i;
                x = SX[a];
                y = SY[a];
                vx = SVX[a];
                vy = SVY[a];
                lold = equal ? lw_equal : l[i] - c;
                if (mon) {
                    const double w = equal ? w_equal : exp(lold);
                    if (w > 0.0) {
                        mx += w * static_cast<double>(x);
                        my += w * static_cast<double>(y);
                    }
                }
            }
            cv_advance<Real, kInit>(m, key, blk, x, y, vx, vy);
            X[i] = x;
            Y[i] = y;
            VX[i] = vx;
            VY[i] = vy;
            const double ll = static_cast<double>(cv_loglik<Real>(m, x, y, ox, oy));
            const double lnew = dead_if_nan(kInit ? ll : lold + ll);
            l[i] = lnew;
            acc = part_merge(acc, part_of(lnew));
        }
    }
    block_partials_at(acc, part_out);
    if (mon) {
        __shared__ double lds[kWaves];
        mx = block_reduce(mx, OpAddD(), lds);
        my = block_reduce(my, OpAddD(), lds);
        if (threadIdx.x == 0) {
            mon_out[0] = mx;
            mon_out[1] = my;
        }
    }
}

// One move over all tiles: from src (through the ancestors anc when the
// last event left them, see cv_tile) into state; src == state moves in place.
// With alt (two buffers, state and alt, pd->cur naming the current one):
// in place in the current buffer, except for a move through the ancestors,
// which reads the current buffer and writes the other (the step's finalize
// then swaps pd->cur). A move in place streams read and written lines of
// the same addresses, measured 20 us (12 %) faster at N = 2^24 than one
// between two buffers.
template <typename Real, bool kInit, int kB>
__global__ void __launch_bounds__(kPB) k_cv_step(Real *state, const Real *src,
    Real *alt, const int *anc, double *__restrict__ l, long long n,
    uint32_t seed, long long t, Real ox, Real oy, CvModelDev m, const PDev *pd,
    Part3 *part, double *mon)
{
    if (alt) {
        const bool gather = anc != nullptr && pd->anc_valid;
        Real *const now = pd->cur ? alt : state;
        src = now;
        state = gather ? (now == alt ? state : alt) : now;
    }
    cv_tile<Real, kInit, kB>(src, src + n, src + 2 * n, src + 3 * n, anc,
        state, state + n, state + 2 * n, state + 3 * n, l, n, blockIdx.x, seed,
        t, ox, oy, m, pd,
        part + blockIdx.x, mon ? mon + 8 * static_cast<long long>(blockIdx.x) : nullptr);
}

// ------------------------------------------------- learned proposal

// Move with the network's Gaussian proposal (FP32 state). Per particle:
// features of (x_{t-1}, y_t), network outputs o (mean offsets o[0..3],
// log-scale offsets o[4..7], clamped to [-5, 3]), the same four normals z
// as the bootstrap move (Philox block t + 1 of the particle's stream),
// u = o_mu + exp(ls) z, x_t = m + s u with the transition mean m and scales
// s, and the weight increment log g + log f - log q
// = ll(x_t) + sum_d (ls_d + z_d^2 / 2 - u_d^2 / 2).
// A block runs chunks [ch0, ch0 + nch) of a 4096-particle tile, 256
// particles (one per thread) per chunk; each wave runs the network on 4
// column tiles of 16 particles of its 64 particles, and the span's weight
// partial goes to out. A filter of more than kLearnedChunkedTiles tiles
// runs whole tiles (nch = kLChunks), its partial the tile partial that the
// merge and the resampling tile prefix read, as for the bootstrap move; a
// smaller one runs one chunk per block (so that a few small filters still
// fill the device), and merge_chunks folds a tile's chunk partials in
// chunk order into the tile partial. The choice depends on the filter's
// size alone, so a filter gives the same results alone and in a batch.
constexpr int kLChunks = kLearnedChunks;
// -ln(2) / 2 rounded up: e^(2 ls) = 0.50007 at the floor
constexpr float kLsFloorHeavy = -0.3465f;
static_assert(kLChunks * kPB == kPTile, "a learned chunk is one block of particles");

__device__ inline Part3 merge_chunks(const Part3 *c)
{
    Part3 acc = c[0];
    for (int k = 1; k < kLChunks; ++k)
        acc = part_merge(acc, c[k]);
    return acc;
}

__device__ inline void cv_learned_span(float *__restrict__ X,
    float *__restrict__ Y, float *__restrict__ VX, float *__restrict__ VY,
    double *__restrict__ l, long long n, long long tile, int ch0, int nch,
    uint32_t seed, long long t, float ox, float oy, const CvModelDev &m,
    const mlp::FwdFrags *__restrict__ F, const PDev *pd, Part3 *out)
{
    __shared__ float fl[kWaves][64][8];
    __shared__ float ol[kWaves][64][8];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int fo = 4 * (lane >> 4);
    const long long base = tile * kPTile;
    const double lw_equal = pd->log_equal;
    const int equal = pd->equal;
    const double c = pd->c;
    const float dt = static_cast<float>(m.dt);
    const float sp = static_cast<float>(m.sd_pos);
    const float sv = static_cast<float>(m.sd_vel);
    const unsigned long long blk = static_cast<unsigned long long>(t) + 1;
    const bool boot = pd->fallback != 0;
    // Under a Student-t likelihood (polynomial tails) the weight p g / q has
    // a finite second moment only if the proposal's variance exceeds half the
    // transition's: e^(2 ls) > 1/2, so ls stays above -ln(2) / 2 there. A
    // Gaussian likelihood keeps it finite for any scale.
    const float ls_min = m.obs_model == 1 ? kLsFloorHeavy : -5.0f;
    Part3 acc{-INFINITY, 0.0, 0.0};
    for (int ch = ch0; ch < ch0 + nch; ++ch) {
        const long long i = base + ch * kPB + wave * 64 + lane;
        const bool valid = i < n;
        float x = 0.0f, y = 0.0f, vx = 0.0f, vy = 0.0f;
        double lold = 0.0;
        if (valid) {
            x = X[i];
            y = Y[i];
            vx = VX[i];
            vy = VY[i];
            lold = equal ? lw_equal : l[i] - c;
        }
        const float mx = x + dt * vx, my = y + dt * vy;
        fl[wave][lane][0] = ox - mx;
        fl[wave][lane][1] = oy - my;
        fl[wave][lane][2] = vx;
        fl[wave][lane][3] = vy;
        fl[wave][lane][4] = ox - x;
        fl[wave][lane][5] = oy - y;
        __syncthreads();
        // after a collapse the filter moves with the bootstrap proposal:
        // zero network outputs (u = z, increment 0)
#pragma unroll 1
        for (int ct = 0; ct < (boot ? 0 : 4); ++ct) {
            const int pp = ct * 16 + (lane & 15);
            float g[6];
#pragma unroll
            for (int k = 0; k < 6; ++k)
                g[k] = fl[wave][pp][k];
            mlp::Fwd r;
            mlp::forward_tile(F, mlp::input_frag(F, g), r);
            if (fo < 8) {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    ol[wave][pp][fo + v] = r.o[v];
            }
        }
        if (boot) {
#pragma unroll
            for (int v = 0; v < 8; ++v)
                ol[wave][lane][v] = 0.0f;
        }
        __syncthreads();
        if (valid) {
            const uint32_t key = seed_key(seed, static_cast<unsigned long long>(i));
            const u32x4 wd = philox_block(key, blk);
            const normal_pair<float> p0 = cv_normals<float>(wd.v[0], wd.v[1]);
            const normal_pair<float> p1 = cv_normals<float>(wd.v[2], wd.v[3]);
            const float z[4] = {p0.z0, p0.z1, p1.z0, p1.z1};
            const float mp[4] = {mx, my, vx, vy};
            const float sc[4] = {sp, sp, sv, sv};
            float s1[4];
            float inc = 0.0f;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                float ls = ol[wave][lane][4 + d];
                ls = ls < ls_min ? ls_min : (ls > 3.0f ? 3.0f : ls);
                const float u = ol[wave][lane][d] + expf(ls) * z[d];
                s1[d] = mp[d] + sc[d] * u;
                inc += ls + 0.5f * z[d] * z[d] - 0.5f * u * u;
            }
            X[i] = s1[0];
            Y[i] = s1[1];
            VX[i] = s1[2];
            VY[i] = s1[3];
            const float ll = cv_loglik<float>(m, s1[0], s1[1], ox, oy);
            // an infinite normal makes inc = inf - inf: the particle dies
            const double lnew = dead_if_nan(lold + static_cast<double>(inc) +
                static_cast<double>(ll));
            l[i] = lnew;
            acc = part_merge(acc, part_of(lnew));
        }
        __syncthreads();
    }
    block_partials_at(acc, out);
}

// The forward pass's fragments copied into LDS by the whole block (the
// forward tiles then read them from LDS rather than through the vector
// cache); the backward-only fragments stay in global memory.
static_assert(sizeof(mlp::FwdFrags) % 16 == 0, "fragments copy in 16-byte units");
static_assert(alignof(mlp::FwdFrags) >= 16, "fragments copy in 16-byte units");
__device__ inline void frags_to_lds(const mlp::FwdFrags *F, mlp::FwdFrags *dst)
{
    const uint4 *s = reinterpret_cast<const uint4 *>(F);
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    for (int k = threadIdx.x; k < static_cast<int>(sizeof(mlp::FwdFrags) / 16); k += kPB)
        d[k] = s[k];
    __syncthreads();
}

// chunked: kLChunks blocks per tile, the partial of block b at out[b];
// otherwise one block per tile
__global__ void __launch_bounds__(kPB) k_cv_learned(float *__restrict__ state,
    double *__restrict__ l, long long n, uint32_t seed, long long t, float ox,
    float oy, CvModelDev m, const mlp::Frags *__restrict__ F, const PDev *pd,
    int chunked, Part3 *out)
{
    __shared__ mlp::FwdFrags sF;
    frags_to_lds(&F->fwd, &sF);
    const long long tile = chunked ? blockIdx.x / kLChunks : blockIdx.x;
    const int ch0 = chunked ? static_cast<int>(blockIdx.x % kLChunks) : 0;
    cv_learned_span(state, state + n, state + 2 * n, state + 3 * n, l, n, tile,
        ch0, chunked ? 1 : kLChunks, seed, t, ox, oy, m, &sF, pd, out + blockIdx.x);
}

__global__ void __launch_bounds__(kPB) k_learned_merge(const Part3 *sub,
    long long ntiles, Part3 *part)
{
    const long long tile = static_cast<long long>(blockIdx.x) * kPB + threadIdx.x;
    if (tile < ntiles)
        part[tile] = merge_chunks(sub + tile * kLChunks);
}

// Draws from the word stream (key, first word) in order: Box-Muller pairs
// (FP32 uniforms) for the training transitions.
struct WordStream {
    uint32_t key;
    unsigned long long word;
    __device__ uint32_t next() { return philox_word(key, word++); }
    __device__ float normal(float &cache, bool &have)
    {
        if (have) {
            have = false;
            return cache;
        }
        const uint32_t a = next();
        const uint32_t b = next();
        const normal_pair<float> p = cv_normals<float>(a, b);
        cache = p.z1;
        have = true;
        return p.z0;
    }
};

// Training transitions from the current cloud (x_{t-1}): x_t ~ f, an
// observation y ~ g(. | x_t) of the model (Gaussian, or Student-t with df
// integer <= 28 as z / sqrt(chi2_df / df)), features of (x_{t-1}, y) and
// the standardised target z = (x_t - m) / s. Stream: key
// seed_key(train_seed, i), 64 words per step from word 64 t.
__global__ void __launch_bounds__(kPB) k_cv_transitions(const float *state,
    long long n, CvModelDev m, uint32_t train_seed, long long t, float *feat,
    float *target)
{
    const long long i = static_cast<long long>(blockIdx.x) * kPB + threadIdx.x;
    if (i >= n)
        return;
    const float x = state[i], y = state[n + i];
    const float vx = state[2 * n + i], vy = state[3 * n + i];
    WordStream ws{seed_key(train_seed, static_cast<unsigned long long>(i)),
        64ULL * static_cast<unsigned long long>(t)};
    float cache = 0.0f;
    bool have = false;
    float z[4];
    for (int d = 0; d < 4; ++d)
        z[d] = ws.normal(cache, have);
    const float dt = static_cast<float>(m.dt);
    const float px = x + dt * vx + static_cast<float>(m.sd_pos) * z[0];
    const float py = y + dt * vy + static_cast<float>(m.sd_pos) * z[1];
    float e[2];
    for (int k = 0; k < 2; ++k) {
        float v = ws.normal(cache, have);
        if (m.obs_model != 2) {
            const int df = static_cast<int>(m.obs_df);
            float chi = 0.0f;
            for (int q = 0; q < df; ++q) {
                const float g = ws.normal(cache, have);
                chi += g * g;
            }
            v = v * sqrtf(static_cast<float>(df) / chi);
        }
        e[k] = static_cast<float>(m.obs_scale) * v;
    }
    const float ox = px + e[0], oy = py + e[1];
    float *f = feat + i * 8;
    f[0] = ox - (x + dt * vx);
    f[1] = oy - (y + dt * vy);
    f[2] = vx;
    f[3] = vy;
    f[4] = ox - x;
    f[5] = oy - y;
    f[6] = 0.0f;
    f[7] = 0.0f;
    for (int d = 0; d < 4; ++d)
        target[i * 4 + d] = z[d];
}

// ------------------------------------------------- batched filters
//
// Many independent filters of the same model in one set of arrays (see
// BatchFilter). A step is four launches on one stream: the persistent
// move over every active filter's tiles, one merge + decision block per
// filter, one resampling workgroup per resampling filter, and one thread
// per filter for the post-event update (which also advances the device
// step counter, so that one captured step replays for every step). Each
// filter runs the same per-tile code as a single filter, so its results
// are those of a single-filter run with the same seed and observations.

// kLearned: the learned moves t >= 1 only (nothing at t = 0), so that the
// kernel's registers are the network's alone; the bootstrap variant runs
// every step, or with init_only the initialisation alone (for a learned
// batch, which launches both variants every step).
template <bool kLearned>
__global__ void __launch_bounds__(kPB) k_batch_move(float *state, double *l,
    long long ld, const BatchFilter *bf, const int *tile_filter,
    long long ntiles, const long long *tstep, const double *obs,
    long long tmax, CvModelDev m, const mlp::Frags *net, const PDev *pd,
    Part3 *part, Part3 *sub, int init_only)
{
    const long long t = *tstep;
    if (kLearned ? t == 0 : (init_only && t != 0))
        return;
    // the learned move has kLChunks units per tile: one chunk each for a
    // chunked filter, else the whole tile in the first unit. Units run
    // chunk-major (u = ch ntiles + tile), so that the tiles' first units
    // are the first blocks, spread over every XCD.
    const int per = kLearned ? kLChunks : 1;
    const mlp::FwdFrags *nf = nullptr;
    if constexpr (kLearned) {
        __shared__ mlp::FwdFrags sF;
        frags_to_lds(&net->fwd, &sF);
        nf = &sF;
    }
    for (long long u = blockIdx.x; u < ntiles * per; u += gridDim.x) {
        const long long tile = u % ntiles;
        const int ch = static_cast<int>(u / ntiles);
        const int f = tile_filter[tile];
        const BatchFilter F = bf[f];
        if (t >= F.T)
            continue;
        const long long o = 2 * (static_cast<long long>(f) * tmax + t);
        const float ox = static_cast<float>(obs[o]);
        const float oy = static_cast<float>(obs[o + 1]);
        float *X = state + F.off;
        double *lf = l + F.off;
        const long long lt = tile - F.tile0;
        if (kLearned && learned_chunked(F.ntiles))
            cv_learned_span(X, X + ld, X + 2 * ld, X + 3 * ld, lf, F.n, lt,
                ch, 1, F.seed, t, ox, oy, m, nf, pd + f, sub + tile * kLChunks + ch);
        else if (kLearned) {
            if (ch == 0)
                cv_learned_span(X, X + ld, X + 2 * ld, X + 3 * ld, lf, F.n, lt,
                    0, kLChunks, F.seed, t, ox, oy, m, nf, pd + f, part + tile);
        } else if (t == 0)
            cv_tile<float, true, 1>(X, X + ld, X + 2 * ld, X + 3 * ld, nullptr,
                X, X + ld, X + 2 * ld, X + 3 * ld, lf, F.n,
                lt, F.seed, t, ox, oy, m, pd + f, part + tile, nullptr);
        else
            cv_tile<float, false, 1>(X, X + ld, X + 2 * ld, X + 3 * ld, nullptr,
                X, X + ld, X + 2 * ld, X + 3 * ld, lf, F.n,
                lt, F.seed, t, ox, oy, m, pd + f, part + tile, nullptr);
    }
}

__global__ void __launch_bounds__(kPB) k_batch_finalize(const BatchFilter *bf,
    const long long *tstep, Part3 *part, const Part3 *sub, PDev *pd,
    ResampleDev *prm, double threshold, int scheme, long long tmax,
    double *logz_hist, double *ess_hist, int *rs_hist, double *tpre, int *hard)
{
    const int f = blockIdx.x;
    // the batched resampler's list of tiles to count exactly starts empty
    if (f == 0 && threadIdx.x == 0)
        hard[0] = 0;
    const long long t = *tstep;
    const BatchFilter F = bf[f];
    if (t >= F.T) {
        if (threadIdx.x == 0)
            prm[f].active = 0;
        return;
    }
    if (sub && t > 0 && learned_chunked(F.ntiles)) {
        // learned move: the tile partials from the chunk partials
        for (long long k = threadIdx.x; k < F.ntiles; k += kPB)
            part[F.tile0 + k] = merge_chunks(sub + (F.tile0 + k) * kLChunks);
        __syncthreads();
    }
    // a learned batch's moves (t >= 1) are guarded against collapse
    weight_merge(part + F.tile0, F.ntiles, F.n, t == 0 ? 0 : 1, pd + f,
        sub != nullptr && t > 0 ? 1 : 0);
    __shared__ int act;
    if (threadIdx.x == 0) {
        PDev *p = pd + f;
        const long long h = static_cast<long long>(f) * tmax + t;
        logz_hist[h] = p->logz;
        ess_hist[h] = p->ess;
        act = decide_event(p, prm + f, F.n, threshold, scheme);
        rs_hist[h] = act;
    }
    __syncthreads();
    if (!act)
        return;
    // the approximate prefix of the normalised weights over the filter's
    // tiles (as k_prepare for one particle set), for the tile summaries
    const PDev *p = pd + f;
    const int equal = p->equal;
    const double lmax = p->lmax;
    const double inv = equal ? 1.0 : 1 / p->sum;
    const double inv_n = 1 / static_cast<double>(F.n);
    __shared__ double lds[kWaves];
    block_scan_chunked(F.ntiles,
        [&](long long b) {
            if (equal) {
                const long long rem = F.n - b * kPTile;
                return static_cast<double>(rem < kPTile ? rem : kPTile) * inv_n;
            }
            const Part3 q = part[F.tile0 + b];
            return q.s > 0.0 ? q.s * exp(q.m - lmax) * inv : 0.0;
        },
        [&](long long b, double ex) { tpre[F.tile0 + b] = ex; }, OpAddD(), 0.0,
        lds);
}

__global__ void k_batch_after(const BatchFilter *bf, PDev *pd,
    ResampleDev *prm, const int *rep, int scheme, int nfilters,
    long long *tstep)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nfilters) {
        const BatchFilter F = bf[f];
        after_event(pd + f, prm + f, F.n, rep + F.off, scheme, 1);
    }
    // every other kernel of this step precedes this one on the stream
    if (f == 0)
        *tstep += 1;
}

int batch_step(const BatchDev &b, int move_blocks, const MlpSet *net,
    hipStream_t st)
{
    const mlp::Frags *nf = net ? static_cast<const mlp::Frags *>(net->frags) : nullptr;
    if (net) {
        // the initialisation (t = 0) of a learned batch: the bootstrap variant
        const long long g = move_blocks < b.ntiles ? move_blocks : b.ntiles;
        hipLaunchKernelGGL(k_batch_move<false>, dim3(g), dim3(kPB), 0, st,
            b.state, b.l, b.ld, b.bf, b.tile_filter, b.ntiles,
            static_cast<const long long *>(b.tstep), b.obs, b.tmax, b.m, nf,
            static_cast<const PDev *>(b.pd), static_cast<Part3 *>(b.part),
            static_cast<Part3 *>(b.lsub), 1);
    }
    hipLaunchKernelGGL(net ? k_batch_move<true> : k_batch_move<false>,
        dim3(move_blocks), dim3(kPB), 0, st,
        b.state, b.l, b.ld, b.bf, b.tile_filter, b.ntiles,
        static_cast<const long long *>(b.tstep), b.obs, b.tmax, b.m, nf,
        static_cast<const PDev *>(b.pd), static_cast<Part3 *>(b.part),
        static_cast<Part3 *>(b.lsub), 0);
    hipLaunchKernelGGL(k_batch_finalize, dim3(b.nfilters), dim3(kPB), 0, st,
        b.bf, static_cast<const long long *>(b.tstep),
        static_cast<Part3 *>(b.part),
        net ? static_cast<const Part3 *>(b.lsub) : nullptr, b.pd, b.prm, b.threshold,
        b.scheme, b.tmax, b.logz_hist, b.ess_hist, b.rs_hist, b.tpre, b.hard);
    VSMCH_HIP(hipGetLastError());
    int rc = batch_resample(b.scheme, b.nfilters, b.bf, b.tile_filter,
        b.ntiles, b.prm, b.l, b.state, b.ld, b.rep, b.tagg, b.toff, b.zdst,
        b.cE, b.csrc, b.tpre, b.segs, b.xstart, b.hard, st);
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL(k_batch_after, dim3((b.nfilters + 255) / 256), dim3(256),
        0, st, b.bf, b.pd, b.prm, static_cast<const int *>(b.rep), b.scheme,
        b.nfilters, b.tstep);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// Normalised weights of one filter of a batch.
int batch_weights(const BatchDev &b, int f, const BatchFilter &F, double *W,
    hipStream_t st)
{
    hipLaunchKernelGGL(k_weight_read, dim3((F.n + 255) / 256), dim3(256), 0,
        st, static_cast<const double *>(b.l + F.off), F.n,
        static_cast<const PDev *>(b.pd + f), W);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// ------------------------------------------------------------ host side

static long long pblocks(long long n)
{
    const long long b = (n + kPTile - 1) / kPTile;
    return b < 1 ? 1 : b;
}

int particle_create(ParticleSet **out, long long n, int dim, int bytes,
    uint32_t seed, hipStream_t st)
{
    if (n < 1 || dim < 1 || dim > 64 || (bytes != 4 && bytes != 8)) {
        set_error("particle_create: invalid configuration");
        return kErrInvalid;
    }
    if (n > 2147483647LL) {
        set_error("particle_create: at most 2^31 - 1 particles (indices are int32)");
        return kErrInvalid;
    }
    ParticleSet *p = new ParticleSet();
    p->n = n;
    p->dim = dim;
    p->bytes = bytes;
    p->seed = seed;
    p->ws.stream = st;
    const long long nb = pblocks(n);
    bool ok = hipMalloc(&p->state, static_cast<size_t>(n) * dim * bytes) == hipSuccess;
    ok = ok && hipMalloc(&p->logw, sizeof(double) * n) == hipSuccess;
    ok = ok && hipMalloc(&p->scratch, sizeof(double) * n) == hipSuccess;
    ok = ok && hipMalloc(&p->rep, sizeof(int) * n) == hipSuccess;
    ok = ok && hipMalloc(&p->part, sizeof(Part3) * nb) == hipSuccess;
    ok = ok && hipMalloc(&p->mpart, sizeof(double) * 8 * nb) == hipSuccess;
    ok = ok && hipMalloc(&p->mout, sizeof(double) * 8) == hipSuccess;
    ok = ok && hipMalloc(&p->pd, sizeof(PDev)) == hipSuccess;
    ok = ok && p->ws.ensure(n);
    if (!ok) {
        particle_destroy(p);
        set_error("particle_create: allocation failed");
        return kErrAlloc;
    }
    PDev h{};
    h.equal = 1;
    h.log_equal = -log(static_cast<double>(n));
    h.c = h.log_equal;
    h.ess = static_cast<double>(n);
    h.rkey = seed_key(seed, static_cast<unsigned long long>(n));
    h.rword = 0;
    if (hipMemcpy(p->pd, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(p->state, 0, static_cast<size_t>(n) * dim * bytes) != hipSuccess ||
        hipMemset(p->logw, 0, sizeof(double) * n) != hipSuccess) {
        particle_destroy(p);
        set_error("particle_create: initialisation failed");
        return kErrHip;
    }
    *out = p;
    return kOk;
}

void particle_destroy(ParticleSet *p)
{
    if (!p)
        return;
    void *ps[] = {p->state, p->logw, p->scratch, p->rep, p->part, p->mpart,
        p->mout, p->pd, p->rng_words, p->cess_buf, p->lsub, p->state_alt,
        p->anc};
    for (void *q : ps)
        if (q)
            (void)hipFree(q);
    if (p->host_active)
        (void)hipHostFree(p->host_active);
    p->ws.release();
    delete p;
}

int weight_update(ParticleSet *p, const double *src, int mode, long long stride)
{
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    if (mode != 2)
        hipLaunchKernelGGL(k_weight_pass, dim3(nb), dim3(kPB), 0, st, p->logw,
            src, p->n, mode, stride, p->pd, static_cast<Part3 *>(p->part));
    // set merges as set_log (the previous weights do not enter the evidence
    // increment), mul as add_log
    const int merge = mode == 3 ? 0 : (mode == 4 ? 1 : mode);
    hipLaunchKernelGGL(k_weight_finalize, dim3(1), dim3(kMergeThreads), 0, st, static_cast<const Part3 *>(p->part),
        nb, p->n, merge, p->pd);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// StateMatrix::copy (state_matrix.hpp:246-253) on the SoA state: particle
// dst takes the pre-copy state of particle idx[dst]. Reads go to the old
// state (gathered into `to`, then copied back), which equals the
// reference's in-order loop whenever every source keeps its own slot
// (idx[src] == src), as in every index of resample_trans_rep_index.
template <typename E>
__global__ void __launch_bounds__(kPB) k_copy_gather(const E *from, E *to,
    const int *idx, long long n, int dim, int *bad)
{
    const long long i = static_cast<long long>(blockIdx.x) * kPB + threadIdx.x;
    if (i >= n)
        return;
    const long long a = idx[i];
    if (a < 0 || a >= n) {
        *bad = 1;   // rejected below: nothing is copied back
        return;
    }
    for (int d = 0; d < dim; ++d)
        to[d * n + i] = from[d * n + a];
}

int state_copy(ParticleSet *p, const int *idx)
{
    hipStream_t st = p->ws.stream;
    const size_t bytes = static_cast<size_t>(p->n) * p->dim * p->bytes;
    if (!p->state_alt)
        VSMCH_HIP(hipMalloc(&p->state_alt, bytes));
    int *bad = reinterpret_cast<int *>(p->mout);
    VSMCH_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
    const dim3 g(static_cast<unsigned int>((p->n + kPB - 1) / kPB));
    if (p->bytes == 4)
        hipLaunchKernelGGL(k_copy_gather<float>, g, dim3(kPB), 0, st,
            static_cast<const float *>(p->state), static_cast<float *>(p->state_alt),
            idx, p->n, p->dim, bad);
    else
        hipLaunchKernelGGL(k_copy_gather<double>, g, dim3(kPB), 0, st,
            static_cast<const double *>(p->state), static_cast<double *>(p->state_alt),
            idx, p->n, p->dim, bad);
    VSMCH_HIP(hipGetLastError());
    int hbad = 0;
    VSMCH_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    VSMCH_HIP(hipStreamSynchronize(st));
    if (hbad) {
        set_error("copy: an index lies outside [0, N)");
        return kErrInvalid;
    }
    VSMCH_HIP(hipMemcpyAsync(p->state, p->state_alt, bytes, hipMemcpyDeviceToDevice, st));
    return kOk;
}

int read_pdev(ParticleSet *p, PDev *h)
{
    VSMCH_HIP(hipMemcpyAsync(h, p->pd, sizeof(PDev), hipMemcpyDeviceToHost,
        p->ws.stream));
    VSMCH_HIP(hipStreamSynchronize(p->ws.stream));
    return kOk;
}

int weight_read(ParticleSet *p, double *W)
{
    hipLaunchKernelGGL(k_weight_read, dim3((p->n + 255) / 256), dim3(256), 0,
        p->ws.stream, p->logw, p->n, p->pd, W);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

static int resample_event(ParticleSet *p, int scheme, bool prep, int *rep_out,
    int *idx_out, bool defer);

// One resampling event. The decision is made on the device; with
// host_decides the host reads it right away (one synchronisation) and a
// step that does not resample launches nothing more. Otherwise the whole
// event is queued and its kernels return early when it is inactive.
int resample_step(ParticleSet *p, int scheme, double threshold, long long t,
    double *ess_hist, int *rs_hist, int *rep_out, int *idx_out,
    bool host_decides)
{
    hipStream_t st = p->ws.stream;
    // systematic / stratified with N >= 2: one kernel prepares the decision,
    // the draw parameters and the tile prefix from the weight partials
    const bool prep = p->n >= 2 && fast_event_scheme(scheme);
    if (prep)
        hipLaunchKernelGGL(k_prepare, dim3(1), dim3(kMergeThreads), 0, st, p->pd,
            p->ws.prm, static_cast<const Part3 *>(p->part), pblocks(p->n),
            p->n, threshold, t, ess_hist, rs_hist, scheme, p->ws.tpre,
            p->ws.fpred);
    else
        hipLaunchKernelGGL(k_decide, dim3(1), dim3(1), 0, st, p->pd, p->ws.prm,
            p->n, threshold, t, ess_hist, rs_hist);
    VSMCH_HIP(hipGetLastError());
    if (host_decides) {
        if (!p->host_active)
            VSMCH_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->host_active),
                sizeof(int), hipHostMallocDefault));
        VSMCH_HIP(hipMemcpyAsync(p->host_active, &p->ws.prm->active, sizeof(int),
            hipMemcpyDeviceToHost, st));
        VSMCH_HIP(hipStreamSynchronize(st));
        if (!*p->host_active)
            return kOk;
    }
    return resample_event(p, scheme, prep, rep_out, idx_out, false);
}

// The event after its decision (made on the device by k_prepare, k_decide
// or k_step_finalize; its kernels return early when it is inactive):
// counts, ancestors, and the in-place state gather, or with `defer` the
// ancestors of every particle in p->anc for the next move, state untouched.
// prep: the decision kernel prepared the draw parameters and tile prefix.
static int resample_event(ParticleSet *p, int scheme, bool prep, int *rep_out,
    int *idx_out, bool defer)
{
    int *rep = rep_out ? rep_out : p->rep;
    int rc = resample_scheme(p->ws, scheme, p->n, p->n, 0, 0, p->logw,
        kWeightLog, rep, true, prep);
    if (rc != kOk)
        return rc;
    // the counting pass left the rep -> index aggregates in ws.tagg (a
    // single particle takes the fill path, which writes none)
    // after_event follows the state gather
    EventEnd end;
    end.pd = p->pd;
    end.rep = rep;
    end.n = p->n;
    end.scheme = scheme;
    end.words = 0;   // counted by the counting pass (or k_words_used)
    end.deferred = defer ? 1 : 0;
    // a prepared event counted with the fast event, which leaves every
    // tile's sources in its own region
    return trans_rep_index(p->ws, p->n, p->n, rep, defer ? p->anc : idx_out,
        defer ? nullptr : p->state, p->dim, p->n, p->bytes, true, false,
        p->n >= 2, end, prep);
}

// r_d = sum_i W_i buf[i * dim + d0 + d] for 8 components at a time.
__global__ void __launch_bounds__(kPB) k_wsum(const double *buf,
    const double *l, long long n, int dim, int d0, const PDev *pd,
    double *part)
{
    const long long base = static_cast<long long>(blockIdx.x) * kPTile;
    const int nd = (dim - d0) < 8 ? (dim - d0) : 8;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < kPItems; ++j) {
        const long long i = base + j * kPB + threadIdx.x;
        if (i < n) {
            const double w = pd->equal ? 1.0 / static_cast<double>(n) :
                                         exp(l[i] - pd->c);
            if (w > 0.0)
                for (int d = 0; d < nd; ++d)
                    acc[d] += w * buf[i * dim + d0 + d];
        }
    }
    __shared__ double lds[kWaves];
    for (int d = 0; d < nd; ++d) {
        const double v = block_reduce(acc[d], OpAddD(), lds);
        if (threadIdx.x == 0)
            part[blockIdx.x * 8 + d] = v;
    }
}

int weighted_sum(ParticleSet *p, int dim, const double *buf, double *out_dev)
{
    if (dim < 1) {
        set_error("weighted_sum: dim must be positive");
        return kErrInvalid;
    }
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    for (int d0 = 0; d0 < dim; d0 += 8) {
        const int nd = (dim - d0) < 8 ? (dim - d0) : 8;
        hipLaunchKernelGGL(k_wsum, dim3(nb), dim3(kPB), 0, st, buf, p->logw,
            p->n, dim, d0, p->pd, p->mpart);
        hipLaunchKernelGGL(k_moments_finalize, dim3(1), dim3(kPB), 0, st,
            p->mpart, nb, nd, out_dev + d0);
    }
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int moments(ParticleSet *p, int c0, int dim, double *out_dev)
{
    if (dim < 1 || dim > 8 || c0 < 0 || c0 + dim > p->dim) {
        set_error("moments: components out of range (at most 8)");
        return kErrInvalid;
    }
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    if (p->bytes == 4)
        hipLaunchKernelGGL(k_moments<float>, dim3(nb), dim3(kPB), 0, st,
            static_cast<const float *>(p->state), p->logw, p->n, c0, dim,
            p->pd, p->mpart);
    else
        hipLaunchKernelGGL(k_moments<double>, dim3(nb), dim3(kPB), 0, st,
            static_cast<const double *>(p->state), p->logw, p->n, c0, dim,
            p->pd, p->mpart);
    hipLaunchKernelGGL(k_moments_finalize, dim3(1), dim3(kPB), 0, st,
        p->mpart, nb, dim, out_dev);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

// Runs of 4 particles whose loads a move thread issues together (1: the
// move is ALU-heavy and more resident waves hide the loads better than
// longer batches; measured on MI355X at N = 2^24). VSMCH_MOVE_BATCH (1, 2
// or 4) overrides it for tuning.
static int move_batch(int bytes)
{
    static const int env = [] {
        const char *v = getenv("VSMCH_MOVE_BATCH");
        const int b = v ? atoi(v) : 0;
        return (b == 1 || b == 2 || b == 4) ? b : 0;
    }();
    (void)bytes;
    return env ? env : 1;
}

template <typename Real, bool kInit>
static void launch_cv(int batch, long long nb, hipStream_t st, Real *s,
    const Real *src, const int *anc, double *l, long long n, uint32_t seed,
    long long t, Real ox, Real oy, const CvModelDev &m, const PDev *pd,
    Part3 *part, double *mon, Real *alt = nullptr)
{
    if (batch == 4)
        hipLaunchKernelGGL((k_cv_step<Real, kInit, 4>), dim3(nb), dim3(kPB),
            0, st, s, src, alt, anc, l, n, seed, t, ox, oy, m, pd, part, mon);
    else if (batch == 2)
        hipLaunchKernelGGL((k_cv_step<Real, kInit, 2>), dim3(nb), dim3(kPB),
            0, st, s, src, alt, anc, l, n, seed, t, ox, oy, m, pd, part, mon);
    else
        hipLaunchKernelGGL((k_cv_step<Real, kInit, 1>), dim3(nb), dim3(kPB),
            0, st, s, src, alt, anc, l, n, seed, t, ox, oy, m, pd, part, mon);
}

int cv_step(ParticleSet *p, const CvModelDev &m, long long t, double ox,
    double oy, bool with_monitor)
{
    if (p->dim != 4) {
        set_error("cv model needs dim == 4");
        return kErrInvalid;
    }
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    double *mon = with_monitor ? p->mpart : nullptr;
    Part3 *part = static_cast<Part3 *>(p->part);
    const int batch = move_batch(p->bytes);
    if (p->bytes == 4) {
        float *s = static_cast<float *>(p->state);
        const float fx = static_cast<float>(ox), fy = static_cast<float>(oy);
        if (t == 0)
            launch_cv<float, true>(batch, nb, st, s, s, nullptr, p->logw, p->n,
                p->seed, t, fx, fy, m, p->pd, part, nullptr);
        else
            launch_cv<float, false>(batch, nb, st, s, s, nullptr, p->logw, p->n,
                p->seed, t, fx, fy, m, p->pd, part, mon);
    } else {
        double *s = static_cast<double *>(p->state);
        if (t == 0)
            launch_cv<double, true>(batch, nb, st, s, s, nullptr, p->logw, p->n,
                p->seed, t, ox, oy, m, p->pd, part, nullptr);
        else
            launch_cv<double, false>(batch, nb, st, s, s, nullptr, p->logw, p->n,
                p->seed, t, ox, oy, m, p->pd, part, mon);
    }
    hipLaunchKernelGGL(k_weight_finalize, dim3(1), dim3(kMergeThreads), 0, st, static_cast<const Part3 *>(p->part),
        nb, p->n, t == 0 ? 0 : 1, p->pd);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

__global__ void k_store_scalar(const PDev *pd, double *logz_hist, long long t)
{
    logz_hist[t] = pd->logz;
}

static int check_learned(const ParticleSet *p, const CvModelDev &m)
{
    if (p->dim != 4 || p->bytes != 4) {
        set_error("learned proposal needs dim == 4 and FP32 state");
        return kErrInvalid;
    }
    if (m.obs_model != 2 && !(m.obs_df >= 1 && m.obs_df <= 28 &&
            m.obs_df == static_cast<double>(static_cast<int>(m.obs_df)))) {
        set_error("learned proposal: Student-t df must be an integer in [1, 28]");
        return kErrInvalid;
    }
    return kOk;
}

int cv_step_learned(ParticleSet *p, const CvModelDev &m, const MlpSet *net,
    long long t, double ox, double oy)
{
    int rc = check_learned(p, m);
    if (rc != kOk)
        return rc;
    if (t < 1) {
        set_error("learned proposal: moves start at t = 1");
        return kErrInvalid;
    }
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    const bool chunked = learned_chunked(nb);
    if (chunked && !p->lsub)
        VSMCH_HIP(hipMalloc(&p->lsub, sizeof(Part3) * kLChunks * nb));
    hipLaunchKernelGGL(k_cv_learned, dim3(chunked ? nb * kLChunks : nb), dim3(kPB),
        0, st, static_cast<float *>(p->state), p->logw, p->n, p->seed, t,
        static_cast<float>(ox), static_cast<float>(oy), m,
        static_cast<const mlp::Frags *>(net->frags), static_cast<const PDev *>(p->pd),
        chunked ? 1 : 0, static_cast<Part3 *>(chunked ? p->lsub : p->part));
    if (chunked)
        hipLaunchKernelGGL(k_learned_merge, dim3((nb + kPB - 1) / kPB), dim3(kPB),
            0, st, static_cast<const Part3 *>(p->lsub), nb,
            static_cast<Part3 *>(p->part));
    hipLaunchKernelGGL(k_weight_finalize, dim3(1), dim3(kMergeThreads), 0, st,
        static_cast<const Part3 *>(p->part), nb, p->n, 1, p->pd, 1);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int cv_transitions(ParticleSet *p, const CvModelDev &m, long long T,
    const double *ox, const double *oy, int scheme, double threshold,
    uint32_t train_seed, float *feat, float *target)
{
    int rc = check_learned(p, m);
    if (rc != kOk)
        return rc;
    hipStream_t st = p->ws.stream;
    const long long g = (p->n + kPB - 1) / kPB;
    for (long long t = 0; t < T; ++t) {
        if (t > 0) {
            const long long row = (t - 1) * p->n;
            hipLaunchKernelGGL(k_cv_transitions, dim3(g), dim3(kPB), 0, st,
                static_cast<const float *>(p->state), p->n, m, train_seed, t,
                feat + row * 8, target + row * 4);
        }
        rc = cv_step(p, m, t, ox[t], oy[t], false);
        if (rc != kOk)
            return rc;
        rc = resample_step(p, scheme, threshold, t, nullptr, nullptr, nullptr,
            nullptr);
        if (rc != kOk)
            return rc;
    }
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

template <typename Real>
static int cv_iterate_impl(ParticleSet *p, const CvModelDev &m, long long t0,
    long long k, const double *ox, const double *oy, int scheme,
    double threshold, double *ess_d, int *rs_d, double *pos_d, double *logz_d)
{
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    if (!p->state_alt)
        VSMCH_HIP(hipMalloc(&p->state_alt,
            static_cast<size_t>(p->n) * p->dim * p->bytes));
    if (!p->anc)
        VSMCH_HIP(hipMalloc(&p->anc, sizeof(int) * p->n));
    Real *const home = static_cast<Real *>(p->state);
    Real *const alt = static_cast<Real *>(p->state_alt);
    // the state starts in the particle set's own buffer; each move runs in
    // place in the current one, a move through the ancestors writes the
    // other (pd->cur on the device)
    hipLaunchKernelGGL(k_own_buffer, dim3(1), dim3(1), 0, st, p->pd);
    const bool prep = p->n >= 2 && fast_event_scheme(scheme);
    const int batch = move_batch(p->bytes);
    Part3 *part = static_cast<Part3 *>(p->part);
    for (long long s = 0; s < k; ++s) {
        const long long t = t0 + s;
        const Real fx = static_cast<Real>(ox[s]), fy = static_cast<Real>(oy[s]);
        // the move of step s also gives the monitor record of step s - 1
        double *mon = (pos_d && s > 0) ? p->mpart : nullptr;
        const int *anc = (t > 0 && s > 0) ? p->anc : nullptr;
        if (t == 0)
            launch_cv<Real, true>(batch, nb, st, home, home, nullptr, p->logw,
                p->n, p->seed, t, fx, fy, m, p->pd, part, nullptr, alt);
        else
            launch_cv<Real, false>(batch, nb, st, home, home, anc, p->logw, p->n,
                p->seed, t, fx, fy, m, p->pd, part, mon, alt);
        hipLaunchKernelGGL(k_step_finalize, dim3(1), dim3(kMergeThreads), 0, st,
            static_cast<const Part3 *>(part), nb, p->n, t == 0 ? 0 : 1, p->pd,
            p->ws.prm, threshold, s, ess_d, rs_d, logz_d,
            static_cast<const double *>(p->mpart),
            mon ? pos_d + 2 * (s - 1) : nullptr, scheme, prep ? 1 : 0,
            p->ws.tpre, p->ws.fpred, anc != nullptr ? 1 : 0);
        VSMCH_HIP(hipGetLastError());
        const int rc = resample_event(p, scheme, prep, nullptr, nullptr, true);
        if (rc != kOk)
            return rc;
    }
    hipLaunchKernelGGL(k_settle<Real>, dim3((p->n + kPB - 1) / kPB), dim3(kPB),
        0, st, static_cast<const Real *>(alt), home,
        static_cast<const int *>(p->anc), static_cast<const PDev *>(p->pd),
        p->n, p->dim);
    VSMCH_HIP(hipGetLastError());
    if (pos_d && k > 0)
        return moments(p, 0, 2, pos_d + 2 * (k - 1));
    return kOk;
}

int cv_iterate(ParticleSet *p, const CvModelDev &m, long long t0, long long k,
    const double *ox, const double *oy, int scheme, double threshold,
    double *ess_d, int *rs_d, double *pos_d, double *logz_d)
{
    if (p->dim != 4) {
        set_error("cv model needs dim == 4");
        return kErrInvalid;
    }
    if (t0 < 0 || k < 0) {
        set_error("cv_iterate: t0 and k must be non-negative");
        return kErrInvalid;
    }
    if (p->bytes == 4)
        return cv_iterate_impl<float>(p, m, t0, k, ox, oy, scheme, threshold,
            ess_d, rs_d, pos_d, logz_d);
    return cv_iterate_impl<double>(p, m, t0, k, ox, oy, scheme, threshold,
        ess_d, rs_d, pos_d, logz_d);
}

int cv_run(ParticleSet *p, const CvModelDev &m, long long T,
    const double *ox, const double *oy, int scheme, double threshold,
    double *ess_d, int *rs_d, double *pos_d, double *logz_d,
    const MlpSet *net)
{
    if (!net)
        return cv_iterate(p, m, 0, T, ox, oy, scheme, threshold, ess_d, rs_d,
            pos_d, logz_d);
    hipStream_t st = p->ws.stream;
    const long long nb = pblocks(p->n);
    int rc;
    if (net) {
        rc = check_learned(p, m);
        if (rc != kOk)
            return rc;
    }
    for (long long t = 0; t < T; ++t) {
        if (net && t > 0) {
            rc = cv_step_learned(p, m, net, t, ox[t], oy[t]);
        } else {
            // the bootstrap move of step t also yields the "pos" record of
            // step t - 1
            rc = cv_step(p, m, t, ox[t], oy[t], !net && pos_d != nullptr && t > 0);
        }
        if (rc != kOk)
            return rc;
        if (!net && pos_d && t > 0)
            hipLaunchKernelGGL(k_moments_finalize, dim3(1), dim3(kPB), 0, st,
                p->mpart, nb, 2, pos_d + 2 * (t - 1));
        if (logz_d)
            hipLaunchKernelGGL(k_store_scalar, dim3(1), dim3(1), 0, st, p->pd,
                logz_d, t);
        rc = resample_step(p, scheme, threshold, t, ess_d, rs_d, nullptr,
            nullptr);
        if (rc != kOk)
            return rc;
        if (net && pos_d && t + 1 < T) {
            rc = moments(p, 0, 2, pos_d + 2 * t);
            if (rc != kOk)
                return rc;
        }
    }
    if (pos_d && T > 0) {
        rc = moments(p, 0, 2, pos_d + 2 * (T - 1));
        if (rc != kOk)
            return rc;
    }
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

} // namespace vsmch
