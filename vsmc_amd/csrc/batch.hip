// batch.hip -- many independent constant-velocity filters on one GPU.
//
// The filters (uneven particle counts and horizons, one seed each) share
// one padded SoA particle array: filter f owns [off_f, off_f + n_f) with
// off_f a multiple of the 4096-particle tile, so every tile belongs to one
// filter and the per-tile code of the single-filter kernels applies as is.
// A step is four launches (particle.hip, batch_step); the run captures one
// step in a HIP graph and replays it for every step (a device counter
// carries the step index), after an autotuner has timed the persistent
// move kernel's grid on the first steps.
//
// The reference runs one Sampler per filter (include/vsmc/core/sampler.hpp
// :443-476, :735-785); there is no batched counterpart to mirror.
#include "particle.hpp"
#include "mlp.hpp"
#include "../../include/vsmch/device_rng.hpp"

#include <math.h>
#include <vector>

namespace vsmch {

struct BatchSet {
    BatchDev d;
    std::vector<BatchFilter> hf;
    hipStream_t stream = nullptr;
    int move_blocks = 0;
    std::vector<double> tune_ms;        // per candidate
    std::vector<int> tune_blocks;
    double last_run_ms = 0;
};

namespace {

constexpr long long kTileP = 4096;

template <typename T>
bool dalloc(T **p, size_t count)
{
    return hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (count ? count : 1)) ==
        hipSuccess;
}

int reset(BatchSet *b)
{
    BatchDev &d = b->d;
    std::vector<PDev> h(d.nfilters);
    for (int f = 0; f < d.nfilters; ++f) {
        const double n = static_cast<double>(b->hf[f].n);
        PDev p{};
        p.equal = 1;
        p.log_equal = -log(n);
        p.c = p.log_equal;
        p.ess = n;
        p.rkey = seed_key(b->hf[f].seed, static_cast<unsigned long long>(b->hf[f].n));
        p.rword = 0;
        h[f] = p;
    }
    const long long zero = 0;
    VSMCH_HIP(hipMemcpyAsync(d.pd, h.data(), sizeof(PDev) * d.nfilters,
        hipMemcpyHostToDevice, b->stream));
    VSMCH_HIP(hipMemcpyAsync(d.tstep, &zero, sizeof(zero), hipMemcpyHostToDevice,
        b->stream));
    VSMCH_HIP(hipMemsetAsync(d.state, 0, sizeof(float) * 4 * d.ld, b->stream));
    VSMCH_HIP(hipMemsetAsync(d.l, 0, sizeof(double) * d.ld, b->stream));
    VSMCH_HIP(hipMemsetAsync(d.logz_hist, 0, sizeof(double) * d.nfilters * d.tmax,
        b->stream));
    VSMCH_HIP(hipMemsetAsync(d.ess_hist, 0, sizeof(double) * d.nfilters * d.tmax,
        b->stream));
    VSMCH_HIP(hipMemsetAsync(d.rs_hist, 0, sizeof(int) * d.nfilters * d.tmax,
        b->stream));
    VSMCH_HIP(hipStreamSynchronize(b->stream));
    return kOk;
}

} // namespace

void batch_destroy(BatchSet *b)
{
    if (!b)
        return;
    if (b->stream)
        (void)hipStreamSynchronize(b->stream);
    BatchDev &d = b->d;
    void *ps[] = {d.bf, d.tile_filter, d.state, d.l, d.part, d.pd, d.prm, d.rep,
        d.zdst, d.cE, d.csrc, d.tagg, d.toff, d.obs, d.tstep, d.logz_hist,
        d.ess_hist, d.rs_hist, d.tpre, d.segs, d.xstart, d.hard, d.lsub};
    for (void *p : ps)
        (void)hipFree(p);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
    delete b;
}

int batch_create(BatchSet **out, int nfilters, const long long *n,
    const long long *T, const uint32_t *seed, const CvModelDev &m, int scheme,
    double threshold)
{
    if (nfilters < 1) {
        set_error("batch: need at least one filter");
        return kErrInvalid;
    }
    if (scheme != kSystematic && scheme != kStratified) {
        set_error("batch: the systematic or stratified scheme");
        return kErrUnsupported;
    }
    BatchSet *b = new BatchSet();
    BatchDev &d = b->d;
    d.nfilters = nfilters;
    d.m = m;
    d.scheme = scheme;
    d.threshold = threshold;
    b->hf.resize(nfilters);
    long long off = 0, tiles = 0, tmax = 0;
    for (int f = 0; f < nfilters; ++f) {
        if (n[f] < 2 || n[f] > (1LL << 30) || T[f] < 1) {
            delete b;
            set_error("batch: every filter needs 2 <= n <= 2^30 and T >= 1");
            return kErrInvalid;
        }
        BatchFilter &F = b->hf[f];
        F.n = n[f];
        F.off = off;
        F.tile0 = tiles;
        F.ntiles = (n[f] + kTileP - 1) / kTileP;
        F.T = T[f];
        F.seed = seed[f];
        F.pad = 0;
        off += F.ntiles * kTileP;
        tiles += F.ntiles;
        tmax = T[f] > tmax ? T[f] : tmax;
    }
    d.ld = off;
    d.ntiles = tiles;
    d.tmax = tmax;
    std::vector<int> tf(tiles);
    for (int f = 0; f < nfilters; ++f)
        for (long long k = 0; k < b->hf[f].ntiles; ++k)
            tf[b->hf[f].tile0 + k] = f;
    const size_t ft = static_cast<size_t>(nfilters) * tmax;
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && dalloc(&d.bf, nfilters) && dalloc(&d.tile_filter, tiles) &&
        dalloc(&d.state, 4 * d.ld) && dalloc(&d.l, d.ld) &&
        dalloc(reinterpret_cast<double **>(&d.part), 3 * tiles) &&
        dalloc(reinterpret_cast<double **>(&d.lsub), 3 * kLearnedChunks * tiles) &&
        dalloc(&d.pd, nfilters) && dalloc(&d.prm, nfilters) &&
        dalloc(&d.rep, d.ld) && dalloc(&d.zdst, d.ld) && dalloc(&d.cE, d.ld) &&
        dalloc(&d.csrc, d.ld) && dalloc(&d.tagg, 3 * tiles) &&
        dalloc(&d.toff, 3 * (tiles + nfilters)) && dalloc(&d.obs, 2 * ft) &&
        dalloc(&d.tpre, tiles) && dalloc(&d.segs, tiles) &&
        dalloc(&d.xstart, tiles) && dalloc(&d.hard, tiles + 1) &&
        dalloc(&d.tstep, 1) && dalloc(&d.logz_hist, ft) &&
        dalloc(&d.ess_hist, ft) && dalloc(&d.rs_hist, ft);
    if (!ok) {
        batch_destroy(b);
        set_error("batch: device allocation failed");
        return kErrAlloc;
    }
    if (hipMemcpy(d.bf, b->hf.data(), sizeof(BatchFilter) * nfilters,
            hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.tile_filter, tf.data(), sizeof(int) * tiles,
            hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d.obs, 0, sizeof(double) * 2 * ft) != hipSuccess ||
        hipMemset(d.prm, 0, sizeof(ResampleDev) * nfilters) != hipSuccess) {
        batch_destroy(b);
        set_error("batch: initialisation failed");
        return kErrHip;
    }
    int rc = reset(b);
    if (rc != kOk) {
        batch_destroy(b);
        return rc;
    }
    *out = b;
    return kOk;
}

int batch_set_obs(BatchSet *b, const double *obs_host)
{
    VSMCH_HIP(hipMemcpy(b->d.obs, obs_host,
        sizeof(double) * 2 * b->d.nfilters * b->d.tmax, hipMemcpyHostToDevice));
    return kOk;
}

// Runs every filter from t = 0 to its horizon. use_graph: capture one step
// in a HIP graph and replay it. The first steps (t >= 1) time each move
// grid candidate once; the fastest is kept for the rest of the run.
int batch_run(BatchSet *b, const MlpSet *net, bool use_graph)
{
    int rc = reset(b);
    if (rc != kOk)
        return rc;
    BatchDev &d = b->d;
    hipStream_t st = b->stream;
    int cus = 256;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipGetDeviceProperties(&prop, dev) == hipSuccess &&
            prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    // candidates: persistent grids of 1, 2, 4 and 8 blocks per CU, capped
    // at one block per unit of the move (a tile; a 256-particle chunk of
    // one for the learned move)
    const long long units = net ? d.ntiles * kLearnedChunks : d.ntiles;
    std::vector<int> cand;
    for (int k : {1, 2, 4, 8}) {
        long long g = static_cast<long long>(cus) * k;
        if (g > units)
            g = units;
        if (cand.empty() || cand.back() != static_cast<int>(g))
            cand.push_back(static_cast<int>(g));
    }
    hipEvent_t e0, e1, r0, r1;
    VSMCH_HIP(hipEventCreate(&e0));
    VSMCH_HIP(hipEventCreate(&e1));
    VSMCH_HIP(hipEventCreate(&r0));
    VSMCH_HIP(hipEventCreate(&r1));
    VSMCH_HIP(hipEventRecord(r0, st));
    long long t = 0;
    int best = cand.back();
    b->tune_ms.assign(cand.size(), 0.0);
    b->tune_blocks = cand;
    // t = 0 (initialisation) with the largest grid
    rc = batch_step(d, best, net, st);
    ++t;
    // autotune on the next steps: each candidate times one real step (the
    // results do not depend on the grid)
    double best_ms = 1e300;
    for (size_t c = 0; c < cand.size() && t < d.tmax && rc == kOk; ++c, ++t) {
        VSMCH_HIP(hipEventRecord(e0, st));
        rc = batch_step(d, cand[c], net, st);
        VSMCH_HIP(hipEventRecord(e1, st));
        VSMCH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        VSMCH_HIP(hipEventElapsedTime(&ms, e0, e1));
        b->tune_ms[c] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = cand[c];
        }
    }
    b->move_blocks = best;
    if (rc == kOk && t < d.tmax) {
        if (use_graph) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            VSMCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int crc = batch_step(d, best, net, st);
            hipGraph_t captured = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &captured);
            if (crc != kOk)
                rc = crc;
            else if (ec != hipSuccess) {
                set_error(std::string("batch: graph capture failed: ") +
                    hipGetErrorString(ec));
                rc = kErrHip;
            }
            g = captured;
            if (rc == kOk) {
                VSMCH_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                for (; t < d.tmax; ++t)
                    VSMCH_HIP(hipGraphLaunch(ge, st));
                VSMCH_HIP(hipGraphExecDestroy(ge));
            }
            if (g)
                (void)hipGraphDestroy(g);
        } else {
            for (; t < d.tmax && rc == kOk; ++t)
                rc = batch_step(d, best, net, st);
        }
    }
    VSMCH_HIP(hipEventRecord(r1, st));
    VSMCH_HIP(hipStreamSynchronize(st));
    float run_ms = 0;
    (void)hipEventElapsedTime(&run_ms, r0, r1);
    b->last_run_ms = run_ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(r0);
    (void)hipEventDestroy(r1);
    return rc;
}

int batch_histories(BatchSet *b, double *logz, double *ess, int *rs)
{
    const BatchDev &d = b->d;
    const size_t ft = static_cast<size_t>(d.nfilters) * d.tmax;
    if (logz)
        VSMCH_HIP(hipMemcpy(logz, d.logz_hist, sizeof(double) * ft,
            hipMemcpyDeviceToHost));
    if (ess)
        VSMCH_HIP(hipMemcpy(ess, d.ess_hist, sizeof(double) * ft,
            hipMemcpyDeviceToHost));
    if (rs)
        VSMCH_HIP(hipMemcpy(rs, d.rs_hist, sizeof(int) * ft,
            hipMemcpyDeviceToHost));
    return kOk;
}

// Filter f's state (4 x n, component stride n) and normalised weights
// (both device, either may be null).
int batch_read_filter(BatchSet *b, int f, float *state, double *W)
{
    if (f < 0 || f >= b->d.nfilters) {
        set_error("batch: filter index out of range");
        return kErrInvalid;
    }
    const BatchFilter &F = b->hf[f];
    if (state)
        VSMCH_HIP(hipMemcpy2DAsync(state, sizeof(float) * F.n,
            b->d.state + F.off, sizeof(float) * b->d.ld, sizeof(float) * F.n, 4,
            hipMemcpyDeviceToDevice, b->stream));
    if (W) {
        const int rc = batch_weights(b->d, f, F, W, b->stream);
        if (rc != kOk)
            return rc;
    }
    VSMCH_HIP(hipStreamSynchronize(b->stream));
    return kOk;
}

void batch_info(const BatchSet *b, long long *tmax, long long *particles,
    int *move_blocks, double *run_ms)
{
    if (tmax)
        *tmax = b->d.tmax;
    if (particles)
        *particles = b->d.ld;
    if (move_blocks)
        *move_blocks = b->move_blocks;
    if (run_ms)
        *run_ms = b->last_run_ms;
}

} // namespace vsmch
