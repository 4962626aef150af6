// particle.hpp -- device particle set shared by particle.hip and capi.hip.
#pragma once

#include "internal.hpp"

namespace vsmch {

// Device-resident weight and stream state of one particle set.
struct PDev {
    double lmax;       // max of l
    double sum;        // sum exp(l - lmax)
    double c;          // log W_i = l_i - c (when !equal)
    double ess;
    double logz;       // accumulated log-evidence
    double logz_inc;
    double log_equal;  // -log N
    unsigned long long rword;  // resampling stream: next unread word
    unsigned int rkey;         // resampling stream key
    long long nresampled;      // resampling events so far
    int equal;                 // weights are 1/N (l ignored)
    int resampled;             // last decision
    // the last event left its ancestors in ParticleSet::anc instead of
    // gathering the state: the next move reads particle i from anc[i]
    int anc_valid;
    // cv_iterate: the buffer that holds the current state (0: the particle
    // set's own, 1: the alternate); a move runs in place unless it gathers
    // through the ancestors, which writes the other buffer
    int cur;
    // learned proposal: a learned step's ESS collapsed below
    // kLearnedCollapse N, and the filter's remaining moves take the bootstrap
    // proposal (cleared by set_log / the initialisation)
    int fallback;
};

// ESS fraction below which a learned step counts as collapsed.
constexpr double kLearnedCollapse = 1e-3;

// After a resampling event: equal weights, advance the resampling stream.
// With `words` the event's stream use is computed here (systematic: the one
// constructor word; stratified: up to and including the first uniform above
// s_{M-2}, transform.hpp:61-69).
__device__ inline void after_event(PDev *pd, ResampleDev *prm, long long n,
    const int *rep, int scheme, int words, int deferred = 0)
{
    if (!prm->active)
        return;
    pd->anc_valid = deferred;
    if (words) {
        unsigned long long used = 1;
        if (scheme == kStratified || scheme == kMultinomial) {
            const long long jend = n - rep[n - 1];
            used = static_cast<unsigned long long>(jend + 1 < n ? jend + 1 : n);
        }
        if (n < 2)
            used = scheme == kSystematic ? 1 : 0;
        prm->words_used = used;
    }
    pd->nresampled += 1;
    pd->rword += prm->words_used;
    pd->equal = 1;
    pd->log_equal = -log(static_cast<double>(n));
    pd->c = pd->log_equal;
    pd->ess = static_cast<double>(n);
}

// Model constants of the constant-velocity PF, precomputed on the host.
struct CvModelDev {
    double dt;
    double sd_pos, sd_vel;
    double sd_pos0, sd_vel0;
    double obs_scale, inv_scale, obs_df, inv_df;
    double ll_const;   // normalising constant added for obs_model 1, 2
    int obs_model;
    int pad;
};

struct ParticleSet {
    long long n = 0;
    int dim = 0;
    int bytes = 0;
    uint32_t seed = 0;
    void *state = nullptr;
    double *logw = nullptr;
    double *scratch = nullptr;
    int *rep = nullptr;
    void *part = nullptr;
    double *mpart = nullptr;
    double *mout = nullptr;
    PDev *pd = nullptr;
    unsigned long long *rng_words = nullptr;  // allocated on first request
    int *host_active = nullptr;  // pinned: the decision read by the host
    void *cess_buf = nullptr;    // conditional-ESS search (temper.hip)
    void *lsub = nullptr;        // learned move: weight partial per chunk
    // cv_iterate: the second state buffer (moves alternate between the two)
    // and the ancestors an event leaves for the next move; both allocated on
    // first use
    void *state_alt = nullptr;
    int *anc = nullptr;
    Workspace ws;
};

int weighted_sum(ParticleSet *p, int dim, const double *buf, double *out_dev);

int particle_create(ParticleSet **out, long long n, int dim, int bytes,
    uint32_t seed, hipStream_t st);
void particle_destroy(ParticleSet *p);
// mode 0 set_log, 1 add_log, 2 set_equal, 3 set (linear), 4 mul (linear);
// src[i * stride]
int weight_update(ParticleSet *p, const double *src, int mode,
    long long stride = 1);
// StateMatrix::copy with a device index array of n entries
int state_copy(ParticleSet *p, const int *idx);
int read_pdev(ParticleSet *p, PDev *h);
int weight_read(ParticleSet *p, double *W);
int resample_step(ParticleSet *p, int scheme, double threshold, long long t,
    double *ess_hist, int *rs_hist, int *rep_out, int *idx_out,
    bool host_decides = false);
int moments(ParticleSet *p, int c0, int dim, double *out_dev);
int cv_step(ParticleSet *p, const CvModelDev &m, long long t, double ox,
    double oy, bool with_monitor);
// Steps t0 .. t0 + k - 1 of the bootstrap filter (Sampler::initialize when
// t0 == 0, then iterate): moves alternate between the particle set's two
// state buffers and read each particle through the ancestors the previous
// event left, so the events gather nothing; at the end the state is
// settled into the set's own buffer. Observations and histories are
// indexed 0 .. k - 1 (device histories, each may be null).
int cv_iterate(ParticleSet *p, const CvModelDev &m, long long t0, long long k,
    const double *ox, const double *oy, int scheme, double threshold,
    double *ess_d, int *rs_d, double *pos_d, double *logz_d);
struct MlpSet;
// net: the learned proposal for the moves t >= 1 (null: bootstrap)
int cv_run(ParticleSet *p, const CvModelDev &m, long long T,
    const double *ox, const double *oy, int scheme, double threshold,
    double *ess_d, int *rs_d, double *pos_d, double *logz_d,
    const MlpSet *net = nullptr);
int cv_step_learned(ParticleSet *p, const CvModelDev &m, const MlpSet *net,
    long long t, double ox, double oy);
// Runs the bootstrap filter over the observations and, before each move
// t >= 1, writes n training transitions from the current cloud:
// feat[(t - 1) n + i][8], target[(t - 1) n + i][4] (device).
int cv_transitions(ParticleSet *p, const CvModelDev &m, long long T,
    const double *ox, const double *oy, int scheme, double threshold,
    uint32_t train_seed, float *feat, float *target);

// 256-particle chunks per 4096-particle tile of the learned move
constexpr int kLearnedChunks = 16;
// A filter of at most this many 4096-particle tiles runs the learned move
// one chunk per workgroup (else one tile per workgroup).
constexpr long long kLearnedChunkedTiles = 2;
__host__ __device__ inline bool learned_chunked(long long ntiles)
{
    return ntiles <= kLearnedChunkedTiles;
}

// Device arrays of a batch of filters (batch.hip owns them).
struct BatchDev {
    int nfilters = 0;
    long long ld = 0;         // padded particle count: component stride
    long long ntiles = 0;
    long long tmax = 0;
    BatchFilter *bf = nullptr;
    int *tile_filter = nullptr;
    float *state = nullptr;   // 4 x ld FP32 (x, y, vx, vy)
    double *l = nullptr;      // ld log-weights
    void *part = nullptr;     // per-tile weight partials
    void *lsub = nullptr;     // learned move: weight partial per 256-particle chunk
    PDev *pd = nullptr;       // per filter
    ResampleDev *prm = nullptr;
    int *rep = nullptr;
    int *zdst = nullptr;
    int *cE = nullptr;
    int *csrc = nullptr;
    long long *tagg = nullptr;   // 3 per tile
    long long *toff = nullptr;   // 3 per (tile + filter)
    double *tpre = nullptr;      // approximate prefix of the weights per tile
    fold::TileSeg *segs = nullptr;  // tile summaries
    double *xstart = nullptr;    // exact start of every tile
    int *hard = nullptr;         // count, then the tiles counted exactly
    double *obs = nullptr;       // [nfilters][tmax][2]
    long long *tstep = nullptr;  // device step counter
    double *logz_hist = nullptr; // [nfilters][tmax]
    double *ess_hist = nullptr;
    int *rs_hist = nullptr;
    CvModelDev m{};
    int scheme = 0;
    double threshold = 0.5;
};

// Enqueues one step of every filter (four launches, five with the learned
// proposal; see particle.hip).
int batch_step(const BatchDev &b, int move_blocks, const MlpSet *net,
    hipStream_t st);
int batch_weights(const BatchDev &b, int f, const BatchFilter &F, double *W,
    hipStream_t st);

// batch.hip
struct BatchSet;
int batch_create(BatchSet **out, int nfilters, const long long *n,
    const long long *T, const uint32_t *seed, const CvModelDev &m, int scheme,
    double threshold);
void batch_destroy(BatchSet *b);
int batch_set_obs(BatchSet *b, const double *obs_host);
int batch_run(BatchSet *b, const MlpSet *net, bool use_graph);
int batch_histories(BatchSet *b, double *logz, double *ess, int *rs);
int batch_read_filter(BatchSet *b, int f, float *state, double *W);
void batch_info(const BatchSet *b, long long *tmax, long long *particles,
    int *move_blocks, double *run_ms);

// gmm.hip -- tempered SMC for a Gaussian mixture
struct GmmSet;
int gmm_create(GmmSet **out, long long n, int c, uint32_t seed, double cess,
    int moves, int rounds, int scheme, double threshold, const double *data,
    long long ndata, hipStream_t st);
void gmm_destroy(GmmSet *s);
ParticleSet *gmm_particles(GmmSet *s);
int gmm_move(GmmSet *s, int kind, long long t, int m);
int gmm_set_alpha(GmmSet *s, double alpha);
int gmm_run(GmmSet *s, long long max_steps, double *alpha_d, double *ess_d,
    int *rs_d, double *logz_d, double *integ_d, double *acc_d,
    long long *steps_out);

// temper.hip -- conditional ESS of an increment d of the log-likelihoods
// ll (device, n doubles) and the largest d in (0, dmax] that keeps it at
// or above target (kCand-point search, `rounds` passes after the first)
int particle_cess(ParticleSet *p, const double *ll, double d, double *out);
int particle_cess_solve(ParticleSet *p, const double *ll, double target,
    double dmax, int rounds, double *out);

// rng.hip
int stream_fill(int kind, unsigned long long key, unsigned long long offset,
    long long n, void *out, int mode, hipStream_t st);
int normal_fill(int kind, unsigned long long key, unsigned long long offset,
    long long n, double mean, double sd, int form, double *out,
    hipStream_t st);

} // namespace vsmch
