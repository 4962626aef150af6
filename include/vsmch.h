/*
 * vsmch.h -- C ABI of libvsmch, the MI355X (gfx950) backend of vSMC's
 * per-step particle path.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer (hipMalloc'd) unless the
 *    parameter name ends in _host. Sizes are plain integers.
 *  - Every function returns a vsmch_status; vsmch_last_error() describes
 *    the last failure. The reference C API returns void
 *    (include/vsmc/vsmc.h:286-310); status codes replace its asserts.
 *  - Work is queued on the library stream (vsmch_set_stream; default: the
 *    null stream). Functions that return host values synchronise it.
 *  - Threads: the library stream, the last error and the scratch buffers of
 *    the free functions (resample transforms, u01 sequences, schemes) belong
 *    to the calling host thread, so those functions may be called from
 *    several host threads at once, as the reference's vsmc_resample_*
 *    (lib/src/vsmc_resample.cpp) can. An object (particle set, batch,
 *    network, genealogy) is used by one host thread at a time.
 *
 * Each entry point cites the reference interface it replaces.
 */
#ifndef VSMCH_H
#define VSMCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int vsmch_status;
#define VSMCH_OK 0
#define VSMCH_ERR_INVALID 1
#define VSMCH_ERR_HIP 2
#define VSMCH_ERR_ALLOC 3
#define VSMCH_ERR_UNSUPPORTED 4
#define VSMCH_ERR_WEIGHTS 5

/* Stream kinds: Philox4x32-10 (include/vsmc/rng/philox.hpp:309) and
 * Threefry4x64-20 (include/vsmc/rng/threefry.hpp:389). */
#define VSMCH_PHILOX4X32 0
#define VSMCH_THREEFRY4X64 1

/* Same order as vsmc_resample_scheme (include/vsmc/vsmc.h:277-284) and
 * vsmc::ResampleScheme (include/vsmc/internal/defines.hpp:49-56). */
typedef enum {
    VSMCH_RESAMPLE_MULTINOMIAL = 0,
    VSMCH_RESAMPLE_STRATIFIED = 1,
    VSMCH_RESAMPLE_SYSTEMATIC = 2,
    VSMCH_RESAMPLE_RESIDUAL = 3,
    VSMCH_RESAMPLE_RESIDUAL_STRATIFIED = 4,
    VSMCH_RESAMPLE_RESIDUAL_SYSTEMATIC = 5
} vsmch_resample_scheme;

/* The calling host thread's last error. */
const char *vsmch_last_error(void);
int vsmch_version(void);
/* Library stream (a hipStream_t) of the calling host thread. */
vsmch_status vsmch_set_stream(void *stream);
vsmch_status vsmch_synchronize(void);

/* ------------------------------------------------------------- streams */

/* Words [word_offset, word_offset + n) of the engine seeded with key0:
 * u32 words for Philox, u64 words for Threefry. Replaces
 * CounterEngine::operator()(n, r) (include/vsmc/rng/counter.hpp:357-390)
 * and vsmc_rng_rand (include/vsmc/vsmc.h:58). */
vsmch_status vsmch_rng_words(int kind, uint64_t key0, uint64_t word_offset,
    size_t n, void *out);

/* U01Distribution<double> over the same words: u * 2^-32 (Philox),
 * (u >> 11) * 2^-53 (Threefry). Replaces u01_distribution
 * (include/vsmc/rng/u01_distribution.hpp:218-233). */
vsmch_status vsmch_u01_fill(int kind, uint64_t key0, uint64_t word_offset,
    size_t n, double *out);

/* Normals. form 0: batch normal_distribution(rng, n, r, mean, sd)
 * (include/vsmc/rng/normal_distribution.hpp:222-267); form 1: n scalar
 * NormalDistribution draws (:199-216). */
vsmch_status vsmch_normal_fill(int kind, uint64_t key0, uint64_t word_offset,
    size_t n, double mean, double sd, int form, double *out);

/* --------------------------------------------------- resample transforms */

/* resample_trans_u01_rep: counts of M weights against N uniforms, exactly
 * as the sequential merge of include/vsmc/resample/transform.hpp:42-71
 * (bit-exact, including its FP64 accumulation order). Replaces
 * vsmc_resample_trans_u01_rep (include/vsmc/vsmc.h:286-287). */
vsmch_status vsmch_resample_trans_u01_rep(int m, int n, const double *weight,
    const double *u01, int *replication);

/* The counts of vsmch_resample_trans_u01_rep without allocating: no
 * counterpart in the reference. For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Its scratch is one file-scope device array (about
 * 33 MiB per device, held until the process ends), so m and n are at most
 * 4,194,304 (VSMCH_ERR_INVALID beyond). Every call on one device shares
 * that array: calls that overlap on one device race. */
vsmch_status vsmch_resample_trans_u01_rep_noalloc(int m, int n,
    const double *weight, const double *u01, int *replication);

/* resample_trans_u01_index (transform.hpp:75-95); replaces
 * vsmc_resample_trans_u01_index (include/vsmc/vsmc.h:288-289). */
vsmch_status vsmch_resample_trans_u01_index(int m, int n,
    const double *weight, const double *u01, int *index);

/* resample_trans_rep_index (transform.hpp:99-134); replaces
 * vsmc_resample_trans_rep_index (include/vsmc/vsmc.h:290-291). */
vsmch_status vsmch_resample_trans_rep_index(int m, int n,
    const int *replication, int *index);

/* resample_trans_index_rep (transform.hpp:138-149); replaces
 * vsmc_resample_trans_index_rep (include/vsmc/vsmc.h:292-293). */
vsmch_status vsmch_resample_trans_index_rep(int m, int n, const int *index,
    int *replication);

/* resample_trans_residual (transform.hpp:154-168); the reference returns
 * R = N - sum(integ) (include/vsmc/vsmc.h:294-295), here *r_host. */
vsmch_status vsmch_resample_trans_residual(int m, int n, const double *weight,
    double *resid, int *integ, int *r_host);

/* Resample<Scheme>()(M, N, rng, W, rep) with the Philox4x32 stream
 * (key0, *word_offset_host) as the resampling engine; on return
 * *word_offset_host has advanced by the words the reference engine would
 * have consumed. Replaces vsmc_resample / vsmc_resample_<scheme>
 * (include/vsmc/vsmc.h:297-310; the reference's vsmc_rng is a build-
 * dependent engine, here the stream is explicit). */
vsmch_status vsmch_resample(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication,
    vsmch_resample_scheme scheme);

/* vsmc_resample_<scheme> (include/vsmc/vsmc.h:297-308): vsmch_resample
 * with the scheme fixed. */
vsmch_status vsmch_resample_multinomial(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication);
vsmch_status vsmch_resample_stratified(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication);
vsmch_status vsmch_resample_systematic(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication);
vsmch_status vsmch_resample_residual(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication);
vsmch_status vsmch_resample_residual_stratified(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication);
vsmch_status vsmch_resample_residual_systematic(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication);

/* U01 sequences of caller uniforms (device arrays of n doubles).
 * vsmc_rng_u01_sorted (include/vsmc/vsmc.h:76-77, u01_sequence.hpp:54-62):
 * the sums of log(u_i) / (n - i) in order, u01seq[i] = 1 - exp(sum); the
 * sums are the exact sequential ones, the logs within an ulp of libm's. */
vsmch_status vsmch_u01_sorted(int n, const double *u01, double *u01seq);
/* vsmc_rng_u01_stratified (include/vsmc/vsmc.h:79-80,
 * u01_sequence.hpp:67-74): u01[i] / n + i / n with delta = 1 / n. */
vsmch_status vsmch_u01_stratified(int n, const double *u01, double *u01seq);
/* vsmc_rng_u01_systematic (include/vsmc/vsmc.h:82-83,
 * u01_sequence.hpp:80-89): u01seq[0] = u01 / n, then repeated additions of
 * 1 / n, bit-equal to the sequential loop. */
vsmch_status vsmch_u01_systematic(int n, double u01, double *u01seq);

/* Test hooks (no reference counterpart): the exact sequential prefix sums
 * out[i] = (((w0 + w1) + w2) + ... + wi) and the U01SequenceSorted values
 * (include/vsmc/rng/u01_sequence.hpp:130-164) of a Philox stream. */
vsmch_status vsmch_exact_cumsum(size_t n, const double *w, double *out);
vsmch_status vsmch_sorted_uniforms(size_t n, uint32_t key0,
    uint64_t word_offset, double *out);

/* -------------------------------------------------------- particle set */

/* One particle system on the device: N particles, an SoA state of `dim`
 * components (FP32 or FP64), FP64 log-weights, the Philox streams of
 * RNGSetVector (particle i: key seed_key(seed, i); resampling engine:
 * seed_key(seed, N), include/vsmc/rng/rng_set.hpp:88-115,
 * include/vsmc/core/particle.hpp:60-67). Replaces Particle<T> + Weight +
 * StateMatrix<ColMajor, Dim, T> for device states. */
typedef struct vsmch_particle vsmch_particle;

typedef struct {
    int64_t n;            /* particles */
    int dim;              /* state components */
    int state_bytes;      /* 4 (float) or 8 (double) */
    uint32_t seed;        /* Seed::instance().set(seed) before construction */
} vsmch_particle_config;

vsmch_status vsmch_particle_create(vsmch_particle **p,
    const vsmch_particle_config *cfg);
vsmch_status vsmch_particle_destroy(vsmch_particle *p);
/* Device pointers of the state (component d at state + d * n) and of the
 * log-weights. */
vsmch_status vsmch_particle_state(vsmch_particle *p, void **state);
vsmch_status vsmch_particle_log_weight(vsmch_particle *p, double **logw);
/* Copies the whole state (dim * n values) to / from a device array laid
 * out the same way (StateMatrix::read_state / state assignment). */
vsmch_status vsmch_particle_read_state(vsmch_particle *p, void *dst);
vsmch_status vsmch_particle_write_state(vsmch_particle *p, const void *src);
/* Key of particle i's stream and of the resampling engine; the resampling
 * engine's next unread word. */
vsmch_status vsmch_particle_keys(vsmch_particle *p, uint32_t *key_first_host,
    uint32_t *resample_key_host, uint64_t *resample_word_host);

/* Weight::set_equal (include/vsmc/core/weight.hpp:99-103). */
vsmch_status vsmch_weight_set_equal(vsmch_particle *p);
/* Weight::set_log (weight.hpp:144-149) from a device array. */
vsmch_status vsmch_weight_set_log(vsmch_particle *p, const double *logw);
/* Weight::add_log (weight.hpp:168-173) from a device array. */
vsmch_status vsmch_weight_add_log(vsmch_particle *p, const double *inc);
/* Weight::set, mul, set_log and add_log with a stride (weight.hpp:105-171)
 * from a device array: src[i * stride] for particle i (stride >= 1). The
 * weights are kept as logarithms, so set / mul take non-negative linear
 * weights (log 0 = -inf gives weight 0). */
#define VSMCH_WEIGHT_SET_LOG 0
#define VSMCH_WEIGHT_ADD_LOG 1
#define VSMCH_WEIGHT_SET 3
#define VSMCH_WEIGHT_MUL 4
vsmch_status vsmch_weight_update(vsmch_particle *p, int op, const double *src,
    int64_t stride);
/* Weight::ess (weight.hpp:76); also the log-evidence accumulated by
 * set_log/add_log (log sum_i W_prev,i exp(inc_i),
 * user_guide/tex/smc.tex:169-182). Synchronises. */
vsmch_status vsmch_weight_ess(vsmch_particle *p, double *ess_host);
vsmch_status vsmch_weight_log_zconst(vsmch_particle *p, double *logz_host);
/* Normalised linear weights W (Weight::data, weight.hpp:81). */
vsmch_status vsmch_weight_read(vsmch_particle *p, double *W);

/* StateMatrix::copy(N, index) (include/vsmc/core/state_matrix.hpp:
 * 246-253) from a device array of n == N ancestor indices: particle i takes
 * the state particle index[i] had before the call (equal to the reference's
 * in-order loop for the indices resample_trans_rep_index produces, where
 * every source keeps its own slot). An index outside [0, N) is rejected
 * (VSMCH_ERR_INVALID, state unchanged). Synchronises the library stream. */
vsmch_status vsmch_particle_copy(vsmch_particle *p, int64_t n, const int *index);

/* Particle::resample(op, threshold) (include/vsmc/core/particle.hpp:
 * 170-197): if ESS < threshold * N, draws counts with `scheme` from the
 * resampling engine, turns them into ancestors (transform.hpp:99-134),
 * gathers the state in place (state_matrix.hpp:356-364) and sets equal
 * weights. *resampled_host (may be NULL) receives the decision.
 * rep_out / idx_out (device, may be NULL) receive counts and ancestors. */
vsmch_status vsmch_particle_resample(vsmch_particle *p, int scheme,
    double threshold, int *resampled_host, int *rep_out, int *idx_out);

/* Per-particle engine positions (next unread word of particle i's stream),
 * device array of n uint64, zero at creation; device functors that draw a
 * data-dependent number of words keep them here (the reference keeps one
 * CounterEngine per particle, include/vsmc/rng/rng_set.hpp:81-122). */
vsmch_status vsmch_particle_rng_words(vsmch_particle *p, uint64_t **words);
/* n doubles of device scratch owned by the particle set (a functor's
 * per-particle log-weight increments between eval_sp and eval_post). */
vsmch_status vsmch_particle_scratch(vsmch_particle *p, double **scratch);
/* Monitor::eval's dgemv (include/vsmc/core/monitor.hpp:263-266):
 * r_d = sum_i W_i buf[i * dim + d] for a device buffer of n x dim values. */
vsmch_status vsmch_weighted_sum(vsmch_particle *p, int dim,
    const double *buf, double *r_host);

/* Number of resampling events so far (the length of the true entries of
 * Sampler::resampled_history, include/vsmc/core/sampler.hpp:702-707). */
vsmch_status vsmch_particle_resample_count(vsmch_particle *p,
    int64_t *count_host);

/* Monitor::eval (include/vsmc/core/monitor.hpp:242-268): r_d = sum_i W_i
 * f_d(x_i) for f the identity on components [c0, c0 + dim). */
vsmch_status vsmch_monitor_state_moments(vsmch_particle *p, int c0, int dim,
    double *r_host);

/* Adaptive tempering. No counterpart in the reference, which anneals on a
 * fixed schedule (example/gmm/include/gmm.hpp:567-604); the search the
 * mixture sampler (vsmch_gmm_*) runs, for any target whose per-particle
 * log-likelihoods ll (device, n doubles; finite or -inf) the caller holds.
 * Conditional ESS of the increment d >= 0 under the current weights:
 * (sum_i W_i e^{d ll_i})^2 / sum_i W_i e^{2 d ll_i}, in (0, 1].
 * Synchronous. */
vsmch_status vsmch_particle_cess(vsmch_particle *p, const double *ll,
    double d, double *cess_host);
/* The largest d in (0, dmax] with CESS(d) >= target (0 < target <= 1):
 * dmax itself when it qualifies, else found to within dmax / 16^rounds
 * (1 <= rounds <= 64) with CESS(d) >= target. Synchronous. */
vsmch_status vsmch_particle_cess_solve(vsmch_particle *p, const double *ll,
    double target, double dmax, int rounds, double *d_host);

/* ------------------------------------------- constant-velocity PF model */

/* The constant-velocity bootstrap PF of user_guide/cpp/pf_seq.cpp with the
 * constants of example/pf/include/pf.hpp (state x, y, vx, vy; Student-t or
 * Gaussian position observations). obs_model 0: Student-t without the
 * normalising constants (the reference's log_likelihood), 1: with them,
 * 2: Gaussian with constants. */
typedef struct {
    double dt;
    double sd_pos, sd_vel;    /* process noise */
    double sd_pos0, sd_vel0;  /* prior */
    int obs_model;
    double obs_scale, obs_df;
} vsmch_cv_model;

/* PFInit::eval_sp for all particles, then Weight::set_log
 * (user_guide/cpp/pf_seq.cpp:70-86). obs_* are host scalars. */
vsmch_status vsmch_cv_init(vsmch_particle *p, const vsmch_cv_model *m,
    double obs_x, double obs_y);
/* PFMove::eval_sp for all particles at step `iter` (>= 1), then
 * Weight::add_log (pf_seq.cpp:106-126), fused into one pass. */
vsmch_status vsmch_cv_move(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t iter, double obs_x, double obs_y);

/* The whole Sampler loop for T steps (initialize + iterate(T - 1),
 * include/vsmc/core/sampler.hpp:443-476, :735-785) for the model above:
 * per step ESS before resampling, resampled flag, monitor "pos" (the
 * weighted mean of x, y after resampling) and the cumulative log-evidence,
 * all kept on the device and copied to the host arrays at the end (each
 * may be NULL). obs_x/obs_y: host arrays of length T. */
vsmch_status vsmch_cv_run(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t T, const double *obs_x_host, const double *obs_y_host,
    int scheme, double threshold, double *ess_host, int *resampled_host,
    double *pos_host, double *logz_host);

/* Steps iter0 .. iter0 + k - 1 of the same loop: Sampler::initialize when
 * iter0 == 0, then iterate (include/vsmc/core/sampler.hpp:443-476,
 * :746-778), queued on the library stream without synchronising. Each
 * move reads every particle through the ancestors of the previous event
 * (two state buffers in turn), so the events gather nothing; the state is
 * in the particle set's own buffer when the call's work completes.
 * obs_x/obs_y: host arrays of length k. Histories (device, length k; pos
 * 2k; each may be NULL): as vsmch_cv_run, indexed from iter0. */
vsmch_status vsmch_cv_iterate(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t iter0, int64_t k, const double *obs_x_host,
    const double *obs_y_host, int scheme, double threshold, double *ess,
    int *resampled, double *pos, double *logz);

/* ------------------------------------------------ learned proposal ----
 * No counterpart in the reference, which has only the bootstrap proposal
 * (the transition, user_guide/cpp/pf_seq.cpp:111-122). A 6 -> 64 -> 64 -> 8
 * tanh network (bfloat16 weights and activations on the matrix cores, FP32
 * accumulation) maps the features
 *   ox - (x + dt vx), oy - (y + dt vy), vx, vy, ox - x, oy - y
 * (normalised as (f - shift) * scale) to a diagonal Gaussian proposal in
 * units of the transition: mean offsets o[0..3], log-scale offsets
 * o[4..7]. The filter stays exact in expectation: each weight is corrected
 * by log g + log f - log q. Parameter vector (FP32, VSMCH_MLP_PARAMS):
 *   W1[64][6] | b1[64] | W2[64][64] | b2[64] | W3[8][64] | b3[8]. */
#define VSMCH_MLP_PARAMS 5128

typedef struct vsmch_mlp vsmch_mlp;

/* New network with all parameters zero (the bootstrap proposal), shift 0,
 * scale 1, on the library stream. */
vsmch_status vsmch_mlp_create(vsmch_mlp **m);
vsmch_status vsmch_mlp_destroy(vsmch_mlp *m);
/* Host arrays; params may be NULL (kept), shift / scale of 6 (NULL: kept).
 * Restarts the optimiser state. */
vsmch_status vsmch_mlp_set_params(vsmch_mlp *m, const float *params_host,
    const float *shift_host, const float *scale_host);
vsmch_status vsmch_mlp_get_params(vsmch_mlp *m, float *params_host);
/* Device arrays: feat[n][8] (6 used) -> out[n][8]. */
vsmch_status vsmch_mlp_forward(vsmch_mlp *m, int64_t n, const float *feat,
    float *out);
/* Mean Gaussian negative log-likelihood of target[n][4] (standardised
 * transitions) and its gradient: grad (device, VSMCH_MLP_PARAMS + 1
 * floats, the loss last); loss_host may be NULL. */
vsmch_status vsmch_mlp_loss_grad(vsmch_mlp *m, int64_t n, const float *feat,
    const float *target, float *grad, float *loss_host);
/* `steps` Adam steps (lr; betas 0.9, 0.999) on batches of `batch` of the n
 * device samples (batch k mod n / batch: samples k + j n / batch); the
 * losses before each step go to loss_host[steps] (may be NULL). */
vsmch_status vsmch_mlp_train(vsmch_mlp *m, int64_t n, const float *feat,
    const float *target, int64_t batch, int64_t steps, float lr,
    float *loss_host);
/* Training data from the bootstrap filter itself: runs vsmch_cv_run's loop
 * (without histories) over T observations and, before each move t >= 1,
 * writes one transition per particle from the current cloud with an
 * observation drawn from the model: feat[(t - 1) n + i][8] and
 * target[(t - 1) n + i][4] (device, (T - 1) n rows). FP32 state; a
 * Student-t model needs an integer df in [1, 28]. */
vsmch_status vsmch_cv_transitions(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t T, const double *obs_x_host, const double *obs_y_host,
    int scheme, double threshold, uint32_t train_seed, float *feat,
    float *target);
/* One move at step iter >= 1 with the learned proposal, then the weight
 * update (the counterpart of vsmch_cv_move). */
vsmch_status vsmch_cv_move_learned(vsmch_particle *p, const vsmch_cv_model *m,
    const vsmch_mlp *net, int64_t iter, double obs_x, double obs_y);
/* ---------------------------------------------------- batched filters --
 * No counterpart in the reference, which runs one Sampler per filter
 * (include/vsmc/core/sampler.hpp:443-476, :735-785). Many independent
 * filters of the model above, each with its own particle count, horizon
 * and seed (particle keys and resampling engine as a vsmch_particle with
 * that seed), advanced together: each step is one persistent move over all
 * active filters, one merge per filter, one resampling workgroup per
 * resampling filter (systematic or stratified, counts bit-exact with the
 * single-filter path) and a per-filter update; a run replays one captured
 * HIP graph per step after timing the move's grid on its first steps.
 * Every filter's results equal those of vsmch_cv_run (FP32 state) with
 * the same seed and observations. Filters advance in their own stream. */
typedef struct vsmch_batch vsmch_batch;

typedef struct {
    int64_t n;        /* particles, >= 2 */
    int64_t T;        /* steps incl. the initial one, >= 1 */
    uint32_t seed;    /* Seed::instance().set(seed) before its Sampler */
} vsmch_batch_filter;

vsmch_status vsmch_batch_create(vsmch_batch **b, int nfilters,
    const vsmch_batch_filter *filters_host, const vsmch_cv_model *m,
    int scheme, double threshold);
vsmch_status vsmch_batch_destroy(vsmch_batch *b);
/* obs_host[f][t][2] (x, y) for t < Tmax = max_f T_f. */
vsmch_status vsmch_batch_set_obs(vsmch_batch *b, const double *obs_host);
/* Runs every filter from t = 0 to its horizon; net: learned proposal for
 * t >= 1 (NULL: bootstrap); use_graph: replay a captured step. */
vsmch_status vsmch_batch_run(vsmch_batch *b, const vsmch_mlp *net,
    int use_graph);
/* [nfilters][Tmax] host arrays (each may be NULL): cumulative log-evidence,
 * ESS before resampling, resampled flags; entries t >= T_f are 0. */
vsmch_status vsmch_batch_histories(vsmch_batch *b, double *logz_host,
    double *ess_host, int *resampled_host);
/* Filter f's final state (device, 4 x n, component stride n) and
 * normalised weights (device, n); either may be NULL. */
vsmch_status vsmch_batch_read_filter(vsmch_batch *b, int f, float *state,
    double *weight);
/* Tmax, padded particle count, the move grid the autotuner chose and the
 * GPU time of the last run (ms); any may be NULL. */
vsmch_status vsmch_batch_info(vsmch_batch *b, int64_t *tmax,
    int64_t *particles, int *move_blocks, double *run_ms);

/* ------------------------------------ tempered SMC, Gaussian mixture --
 * The sampler of example/gmm/include/gmm.hpp (gmm_state, gmm_init,
 * gmm_move_smc, gmm_move_mu / _lambda / _weight, gmm_path_integrand) on the
 * device, with adaptive tempering: each step takes the largest increment of
 * alpha whose conditional ESS reaches `cess` (the reference anneals on a
 * fixed schedule, gmm.hpp:567-604), reweights by add_log(increment x log
 * likelihood), resamples when ESS < threshold N, then runs `moves`
 * random-walk Metropolis-Hastings moves cycling mu, lambda, weight.
 * Particles: 3c + 2 FP32 components (mu[c], lambda[c], weight[c], log
 * prior, log likelihood) of an internal particle set. */
typedef struct vsmch_gmm vsmch_gmm;

typedef struct {
    int64_t n;            /* particles */
    int ncomp;            /* mixture components c, 2..8 */
    uint32_t seed;        /* Seed::instance().set(seed) */
    double cess;          /* conditional ESS target in (0, 1), e.g. 0.9 */
    int moves;            /* MH moves per step */
    int search_rounds;    /* rounds of the 16-candidate increment search */
    int scheme;           /* stratified or systematic */
    double threshold;     /* resample when ESS < threshold N */
} vsmch_gmm_config;

/* data_host: ndata observations (2..8192). */
vsmch_status vsmch_gmm_create(vsmch_gmm **g, const vsmch_gmm_config *cfg,
    const double *data_host, int64_t ndata);
vsmch_status vsmch_gmm_destroy(vsmch_gmm *g);
/* Initialisation at alpha = 0, then steps until alpha = 1, at most
 * max_steps history entries (the initial one included). Host histories of
 * max_steps entries (any may be NULL): alpha, ESS before resampling,
 * resampled flags, cumulative log-evidence, the path-sampling integrand
 * sum_i W_i loglik_i after the moves, and acceptance rates (3 per step:
 * mu, lambda, weight; -1 where none was proposed). *steps_host: entries. */
vsmch_status vsmch_gmm_run(vsmch_gmm *g, int64_t max_steps,
    int64_t *steps_host, double *alpha_host, double *ess_host,
    int *resampled_host, double *logz_host, double *integrand_host,
    double *accept_host);
/* The sampler's particle set (borrowed: valid until vsmch_gmm_destroy; not
 * to be destroyed by the caller). */
vsmch_status vsmch_gmm_particle(vsmch_gmm *g, vsmch_particle **p);
/* One MH move of kind 0 (mu), 1 (lambda) or 2 (weight) over all particles
 * at temperature alpha, with the draws of step iter, move index m. */
vsmch_status vsmch_gmm_move(vsmch_gmm *g, int kind, double alpha,
    int64_t iter, int m);

/* vsmch_cv_run with the learned proposal for every move t >= 1. */
vsmch_status vsmch_cv_run_learned(vsmch_particle *p, const vsmch_cv_model *m,
    const vsmch_mlp *net, int64_t T, const double *obs_x_host,
    const double *obs_y_host, int scheme, double threshold, double *ess_host,
    int *resampled_host, double *pos_host, double *logz_host);

/* ---------------------------------------------------------------------
 * Ancestry history: ResampleIndex (include/vsmc/resample/index.hpp:46-206)
 * on the device. Row t holds the ancestor index of every particle at
 * iteration t; `idx` arguments are device arrays of n ints (NULL: the
 * identity, a step without resampling), e.g. the ancestors that
 * vsmch_particle_resample writes to idx_out. An ancestor outside [0, n)
 * makes index / index_matrix fail with VSMCH_ERR_INVALID. */
#define VSMCH_ROW_MAJOR 101   /* vsmc::RowMajor (internal/defines.hpp:49) */
#define VSMCH_COL_MAJOR 102   /* vsmc::ColMajor */

typedef struct vsmch_rindex vsmch_rindex;

/* ResampleIndex(N) (index.hpp:53-57) */
vsmch_status vsmch_rindex_create(vsmch_rindex **r, int64_t n);
vsmch_status vsmch_rindex_destroy(vsmch_rindex *r);
/* iter_size() (:62) */
vsmch_status vsmch_rindex_iter_size(vsmch_rindex *r, int64_t *iters);
/* reset() (:65): no iterations, memory kept */
vsmch_status vsmch_rindex_reset(vsmch_rindex *r);
/* clear() (:68): releases the rows */
vsmch_status vsmch_rindex_clear(vsmch_rindex *r);
/* push_back() / push_back(first) (:70-87) */
vsmch_status vsmch_rindex_push_back(vsmch_rindex *r, const int *idx);
/* insert() / insert(first) / insert(iter, first) (:89-119); iter < 0: the
 * last iteration */
vsmch_status vsmch_rindex_insert(vsmch_rindex *r, int64_t iter, const int *idx);
/* index(id, iter) (:121-134): the ancestor at iteration iter of particle id
 * of the last iteration */
vsmch_status vsmch_rindex_index(vsmch_rindex *r, int64_t id, int64_t iter,
    int64_t *idx_host);
/* index_matrix<Layout> / read_index_matrix (:136-148, :156-204): out is a
 * device array of n * iter_size ints; RowMajor: row i is particle i's
 * lineage over the iterations, ColMajor: row t is iteration t. */
vsmch_status vsmch_rindex_index_matrix(vsmch_rindex *r, int layout, int *out);

#ifdef __cplusplus
}
#endif

#endif /* VSMCH_H */
