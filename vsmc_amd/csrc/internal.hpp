// internal.hpp -- host-side API between the C ABI (capi.hip) and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#include "fold.hpp"

namespace vsmch {

// Device-resident parameters of one resampling event. Kernels read them
// through a pointer so that a captured step needs no host round trip.
struct ResampleDev {
    double lmax;          // weights from log-weights: W = exp(l - lmax) * inv_sum
    double inv_sum;
    double inv_resid;     // 1 / exact sum of residuals
    double u0;            // systematic start U * delta
    double delta;         // fl(1 / ndraw)
    double resid_total;   // exact sequential sum of residuals
    long long ndraw;      // number of uniforms (N, or R for residual schemes)
    long long nsample;    // N (residual schemes add integer parts)
    long long M;          // number of weights
    long long integ_sum;  // sum of integer parts (residual schemes)
    unsigned long long word0;  // resampling-stream word at the event start
    unsigned long long words_used;  // words the event consumed
    unsigned int key;     // resampling-stream key
    int active;           // 0: step did not resample, kernels return early
    int err;              // input validation flags
    int equal;            // weights are 1/N (log-weights ignored)
    int nonmono;          // the sorted draws' epilogue decreased somewhere
    int pad;
};

// One filter of a batch: its particles occupy [off, off + n) of the batch's
// arrays (off a multiple of the 4096-particle tile, so no tile holds two
// filters), tiles [tile0, tile0 + ntiles); horizon T steps; particle i has
// the stream key seed_key(seed, i), the resampling engine seed_key(seed, n).
struct BatchFilter {
    long long n;
    long long off;
    long long tile0;
    long long ntiles;
    long long T;
    unsigned int seed;
    int pad;
};

struct PDev;

// A particle set's update when its resampling event completes, launched
// after the state gather (pd == nullptr: none).
struct EventEnd {
    PDev *pd = nullptr;
    const int *rep = nullptr;
    long long n = 0;
    int scheme = 0;
    int words = 0;
    int deferred = 0;   // ancestors written for the next move, state not gathered
};

// Grow-only device workspace shared by the resampling passes.
struct Workspace {
    hipStream_t stream = nullptr;
    long long cap_n = 0;
    long long cap_tiles = 0;
    double *tsum = nullptr;      // per-tile approximate sums
    double *tpre = nullptr;      // approximate exclusive tile prefix
    fold::TileSeg *segs = nullptr;
    double *xstart = nullptr;    // exact tile starts (+ total at [ntiles])
    double *xgstart = nullptr;   // exact group starts
    int *gmode = nullptr;        // 1: group resolved through its group map
    struct GroupSeg *groups = nullptr;
    long long *tagg = nullptr;   // per-tile (zeros, extras, sources) of rep
    long long *toff = nullptr;   // exclusive offsets of tagg (+ totals)
    int *zdst = nullptr;         // childless slots in dst order
    int *cE = nullptr;           // extra-copy prefix of sources with extras
    int *csrc = nullptr;         // those sources
    int *tq = nullptr;           // per tile: slice [tq[2t], tq[2t+1]) of csrc its childless slots take from
    double *v = nullptr;         // materialised uniforms (running max)
    double *aux = nullptr;       // scratch N doubles
    int *rep = nullptr;          // scratch counts
    int *hard = nullptr;         // batched filters: [0] count, then tiles to fold in full
    struct FeSum *fsum = nullptr;  // dense per-tile summaries (particle-set events)
    struct FeCross *fcross = nullptr;  // their one-crossing remainders
    int *fpred = nullptr;        // [0] count, then the tiles a particle set's
                                 // finalize predicts to cross a binade
    int *integ_dummy = nullptr;
    ResampleDev *prm = nullptr;
    long long *scal = nullptr;   // small device scalars
    // fixed: the buffers are views into memory the workspace does not own
    // (a caller's or a file-scope device array); ensure() then never frees
    // or allocates, it only checks that n fits cap_n / cap_tiles
    bool fixed = false;
    std::string err;

    bool ensure(long long n);
    void release();
    bool all_allocated() const;
};

// The calling host thread's workspace on `device` (created on first use).
// Each thread has its own, so concurrent calls from several host threads
// never share buffers; a thread's workspace returns to a per-device pool
// when the thread exits and the next new thread takes it over, after a
// device synchronisation (work the old thread queued may still read it).
Workspace &default_workspace(int device);

// The library stream of the calling host thread (vsmch_set_stream).
hipStream_t library_stream();

// Tiles the fold passes hold a summary for, for n weights (one spare).
inline long long workspace_tiles(long long n) { return (n + kTile - 1) / kTile + 1; }

// GroupSeg: 64 consecutive tiles that share one binade and have no crossing.
struct GroupSeg {
    int flat;
    int e;
    fold::FMap map;
};

constexpr int kGroup = 64;

// A particle set's event summarises first the tiles its finalize predicts
// to cross a binade (fold::predicted_crossing), at most this many.
constexpr int kFePred = 64;

// After a block has written the approximate tile prefix tpre[0, nb): the
// tiles predicted to cross a binade (fold::predicted_crossing) into fpred
// ([0] their count, which may exceed the kFePred entries), in any order.
__device__ inline void list_predicted(const double *tpre, long long nb, int *fpred,
    int *cnt_lds)
{
    if (threadIdx.x == 0)
        *cnt_lds = 0;
    __syncthreads();   // tpre complete for the block
    for (long long t = threadIdx.x; t < nb; t += blockDim.x) {
        if (fold::predicted_crossing(tpre, t, nb)) {
            const int k = atomicAdd(cnt_lds, 1);
            if (k < kFePred)
                fpred[1 + k] = static_cast<int>(t);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0)
        fpred[0] = *cnt_lds;
}

// Status codes (mirrors vsmch_status in include/vsmch.h)
enum {
    kOk = 0,
    kErrInvalid = 1,
    kErrHip = 2,
    kErrAlloc = 3,
    kErrUnsupported = 4,
    kErrWeights = 5,
};

void set_error(const std::string &msg);
const char *get_error();

#define VSMCH_HIP(expr)                                                        \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::vsmch::set_error(std::string(#expr) + ": " +                     \
                hipGetErrorString(e_));                                        \
            return ::vsmch::kErrHip;                                           \
        }                                                                      \
    } while (0)

// ---- host entry points implemented in resample.hip

// Schemes, enum order of include/vsmc/internal/defines.hpp (ResampleScheme)
enum Scheme {
    kMultinomial = 0,
    kStratified = 1,
    kSystematic = 2,
    kResidual = 3,
    kResidualStratified = 4,
    kResidualSystematic = 5,
};

// Weight source kinds for the resampling passes.
enum WeightKind {
    kWeightArray = 0,    // W given (linear, non-negative)
    kWeightLog = 1,      // W = exp(l - lmax) * inv_sum from prm
};

// Schemes whose particle-set events run the fast event (fast_event.hpp):
// the finalize prepares their draw parameters and the weights' tile prefix
// (residual schemes then set their own from the residuals).
inline bool fast_event_scheme(int scheme)
{
    return scheme >= kMultinomial && scheme <= kResidualSystematic;
}


// Replication counts of the transform with explicit uniforms:
// resample_trans_u01_rep(M, N, W, u01, rep) semantics. Device pointers.
int trans_u01_rep(Workspace &ws, long long M, long long N, const double *W,
    const double *u01, int *rep);

// rep -> index (resample_trans_rep_index semantics); optional state gather
// applied in place on an SoA state of `dim` components of elem_bytes each,
// component stride `ld` elements. idx may be null when only gathering.
// force_expand: lay out the expanded list (src repeated rep[src] times)
// even when M == N (resample_trans_u01_index). agg_ready: the per-tile
// aggregates in ws.tagg were produced by the counting pass that wrote rep.
// tile_sources: (M == N, agg_ready) the counting pass also left every
// tile's sources in its own region of ws.csrc / ws.cE (the fast event).
int trans_rep_index(Workspace &ws, long long M, long long N, const int *rep,
    int *idx, void *state, int dim, long long ld, int elem_bytes,
    bool device_active, bool force_expand = false, bool agg_ready = false,
    const EventEnd &end = EventEnd(), bool tile_sources = false);

int trans_index_rep(Workspace &ws, long long M, long long N, const int *idx,
    int *rep);

int trans_residual(Workspace &ws, long long M, long long N, const double *W,
    double *resid, int *integ, long long *R_out);

// Full scheme with the resampling stream (key, word offset): counts for M
// weights and N draws. weight_kind selects the weight source; for
// kWeightLog the caller has set prm->lmax / inv_sum on the device and
// `weights` points at the log-weights. When device_active is true the
// kernels honour prm->active (set on the device) and the host does not
// synchronise; otherwise the host sets everything.
// prepared: (device_active, systematic / stratified) the draw parameters
// and ws.tpre were set by the weight side (k_prepare), so the tile-sum and
// parameter kernels are skipped.
int resample_scheme(Workspace &ws, int scheme, long long M, long long N,
    unsigned int key, unsigned long long word0, const double *weights,
    int weight_kind, int *rep, bool device_active, bool prepared = false);

// Words the last resample_scheme call consumed (read back to the host).
int read_words_used(Workspace &ws, unsigned long long *out);

// Sorted uniforms of the multinomial scheme (U01SequenceSorted) for
// ndraw draws from (key, word0): out[j] (device). For tests.
int sorted_uniforms(Workspace &ws, long long ndraw, unsigned int key,
    unsigned long long word0, double *out);

// Exact sequential prefix sums (for tests): out[i] = fl(...fl(w0 + w1)...)
int exact_cumsum(Workspace &ws, long long n, const double *w, double *out);
// u01_sorted / u01_stratified / u01_systematic over caller uniforms
int u01_sorted_seq(Workspace &ws, long long n, const double *u, double *out);
int u01_stratified_seq(Workspace &ws, long long n, const double *u, double *out);
int u01_systematic_seq(Workspace &ws, long long n, double u, double *out);

// One resampling event for every filter of a batch whose prm[f].active is
// set (tile summaries, counts, scatter and gather: one workgroup per tile
// of the batch, tile_filter giving each tile's filter, ntiles in all; the
// walk over the summaries and the scan: one per filter; tpre: the
// approximate prefix per tile, segs / xstart: scratch of one per tile,
// hard: a list of ntiles + 1 whose count k_batch_finalize clears);
// weights from the log-weights l with prm[f].lmax / inv_sum / equal; FP32
// SoA state with component stride ld; scratch rep / zdst / cE / csrc of
// the batch's particle count, tagg of 3 per tile, toff of 3 per (tile +
// filter).
int batch_resample(int scheme, int nfilters, const BatchFilter *bf,
    const int *tile_filter, long long ntiles, const ResampleDev *prm,
    const double *l, float *state, long long ld, int *rep, long long *tagg,
    long long *toff, int *zdst, int *cE, int *csrc, const double *tpre,
    fold::TileSeg *segs, double *xstart, int *hard, hipStream_t st);

} // namespace vsmch
