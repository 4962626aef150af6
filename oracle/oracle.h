/*
 * oracle.h -- CPU restatement of vSMC's per-step particle path (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This library is the checker the HIP path is compared against. Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library (libvsmch.so) never links or calls it.
 *
 * Every function restates one piece of the reference (vSMC v2.2.0) and cites
 * the file:line it follows. Integer types: counts and indices are int64_t
 * (the reference uses size_t / int); weights and uniforms are double.
 *
 * Parity status of each function (how it is pinned) is recorded in the
 * header comment of oracle.c.
 */
#ifndef VSMC_ORACLE_H
#define VSMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNG -- */

/* Philox4x32-10 block function: 10 rounds, 9 key bumps.
 * include/vsmc/rng/philox.hpp:162-232, :263-269 */
void vo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
    uint32_t out[4]);

/* Threefry4x64-20 block function.
 * include/vsmc/rng/threefry.hpp:148-298, :304-367 */
void vo_threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4],
    uint64_t out[4]);

/* CounterEngine<PhiloxGenerator<u32,4,10>> state.
 * include/vsmc/rng/counter.hpp:288-496 */
typedef struct {
    uint32_t buf[4];
    uint32_t ctr[4];
    uint32_t key[2];
    uint32_t index;
} vo_philox;

typedef struct {
    uint64_t buf[4];
    uint64_t ctr[4];
    uint64_t key[4];
    uint32_t index;
} vo_threefry;

void vo_philox_seed(vo_philox *e, uint32_t s);
uint32_t vo_philox_next(vo_philox *e);
void vo_philox_fill(vo_philox *e, size_t n, uint32_t *r);
void vo_philox_discard(vo_philox *e, uint32_t nskip);
void vo_philox_set_ctr(vo_philox *e, const uint32_t ctr[4]);

void vo_threefry_seed(vo_threefry *e, uint64_t s);
uint64_t vo_threefry_next(vo_threefry *e);
void vo_threefry_fill(vo_threefry *e, size_t n, uint64_t *r);
void vo_threefry_discard(vo_threefry *e, uint64_t nskip);
void vo_threefry_set_ctr(vo_threefry *e, const uint64_t ctr[4]);

/* Stream words of the engine seeded with key0, skipping `offset` words. */
void vo_philox_words(uint32_t key0, uint64_t offset, size_t n, uint32_t *out);
void vo_threefry_words(uint64_t key0, uint64_t offset, size_t n,
    uint64_t *out);

/* U01 [0,1) conversions. include/vsmc/rng/u01.hpp:126-147 */
double vo_u01_co_u32d(uint32_t u);
double vo_u01_co_u64d(uint64_t u);
float vo_u01_co_u32f(uint32_t u);

/* RNG kinds */
#define VO_PHILOX4X32 0
#define VO_THREEFRY4X64 1

/* U01Distribution<double> over a stream: out[k] = u01(word(offset + k)). */
void vo_u01_fill(int kind, uint64_t key0, uint64_t offset, size_t n,
    double *out);

/* Batch normal_distribution(rng, n, r, mean, sd) for n >= 100
 * (include/vsmc/rng/normal_distribution.hpp:222-267); for n < 100 the
 * reference uses the scalar form, which vo_normal_fill_scalar restates. */
void vo_normal_fill(int kind, uint64_t key0, uint64_t offset, size_t n,
    double mean, double sd, double *out);
void vo_normal_fill_scalar(int kind, uint64_t key0, uint64_t offset,
    size_t n, double mean, double sd, double *out);

/* NormalDistribution<RealType> scalar Box-Muller with cached sine half.
 * include/vsmc/rng/normal_distribution.hpp:199-216 */
typedef struct {
    double mean;
    double sd;
    double v;
    int saved;
} vo_normal;
void vo_normal_init(vo_normal *d, double mean, double sd);
double vo_normal_philox(vo_normal *d, vo_philox *e);

/* SeedGenerator<ID, unsigned> (include/vsmc/rng/seed.hpp:84-228) */
typedef struct {
    uint32_t seed;
    uint32_t seed_max;
    uint32_t divisor;
    uint32_t remainder;
} vo_seedgen;
void vo_seed_init(vo_seedgen *g);
void vo_seed_set(vo_seedgen *g, uint32_t s);
uint32_t vo_seed_get(vo_seedgen *g);

/* ------------------------------------------------------------ Weight -- */

/* Weight::set_log / add_log / set_equal on linear normalised weights W.
 * include/vsmc/core/weight.hpp:99-103, :144-173, :200-229.
 * Each returns the ESS; *logz_inc (if not NULL) receives the log-evidence
 * increment log sum_i W_prev,i * exp(inc_i) (user_guide/tex/smc.tex:169-182),
 * computed from the same pass (max + log of the sequential sum). */
double vo_weight_set_equal(size_t n, double *W);
double vo_weight_set_log(size_t n, double *W, const double *logw,
    double *logz_inc);
double vo_weight_add_log(size_t n, double *W, const double *inc,
    double *logz_inc);
double vo_weight_ess(size_t n, const double *W);

/* ------------------------------------------------------------ Resample -- */

/* u01 sequence free functions. include/vsmc/rng/u01_sequence.hpp:53-89 */
void vo_u01_sorted(size_t N, const double *u01, double *out);
void vo_u01_stratified(size_t N, const double *u01, double *out);
void vo_u01_systematic(size_t N, double u01, double *out);

/* resample_trans_* (include/vsmc/resample/transform.hpp:42-168) */
void vo_trans_u01_rep(size_t M, size_t N, const double *W, const double *u,
    int64_t *rep);
void vo_trans_u01_index(size_t M, size_t N, const double *W, const double *u,
    int64_t *idx);
void vo_trans_rep_index(size_t M, size_t N, const int64_t *rep,
    int64_t *idx);
void vo_trans_index_rep(size_t M, size_t N, const int64_t *idx,
    int64_t *rep);
size_t vo_trans_residual(size_t M, size_t N, const double *W, double *resid,
    int64_t *integ);

/* ResampleIndex (include/vsmc/resample/index.hpp:46-206) over a history of
 * T rows of N ancestors (row t at hist + t N): index(id, iter) and
 * index_matrix<RowMajor> (layout 101: out[i T + t]) / <ColMajor> (102:
 * out[t N + i]). */
int64_t vo_rindex_index(size_t N, size_t T, const int64_t *hist, size_t id,
    size_t iter);
void vo_rindex_matrix(size_t N, size_t T, const int64_t *hist, int layout,
    int64_t *out);

/* Resample schemes, enum order of include/vsmc/internal/defines.hpp
 * (ResampleScheme) and include/vsmc/vsmc.h:277-284 */
#define VO_MULTINOMIAL 0
#define VO_STRATIFIED 1
#define VO_SYSTEMATIC 2
#define VO_RESIDUAL 3
#define VO_RESIDUAL_STRATIFIED 4
#define VO_RESIDUAL_SYSTEMATIC 5

/* Resample<Scheme>()(M, N, rng, W, rep) with the lazily drawn u01 sequence
 * of include/vsmc/rng/u01_sequence.hpp:130-259. Returns the number of
 * engine words consumed. */
uint64_t vo_resample(int scheme, size_t M, size_t N, vo_philox *rng,
    const double *W, int64_t *rep);

/* The u01 sequence a scheme's transform reads (the first `nu` entries,
 * as the lazy sequence would produce them), for tests. */
void vo_scheme_u01(int scheme, size_t N, vo_philox *rng, size_t nu,
    double *out);

/* Sequential left fold of w (transform.hpp:61-63). */
void vo_cumsum(size_t n, const double *w, double *s);

/* StateMatrix<RowMajor|ColMajor, dim, double>::copy
 * (include/vsmc/core/state_matrix.hpp:246-262, :356-364) */
void vo_copy_rowmajor(size_t N, size_t dim, double *x, const int64_t *idx);
void vo_copy_colmajor(size_t N, size_t dim, double *x, const int64_t *idx);

/* Monitor::eval dgemv: r_d = sum_i W_i buf[i*dim + d]
 * (include/vsmc/core/monitor.hpp:258-266) */
void vo_monitor_moments(size_t N, size_t dim, const double *buf,
    const double *W, double *r);

/* ------------------------------------------------------ PF (model) -- */

typedef struct {
    int64_t N;            /* particles */
    int64_t T;            /* steps incl. the initial one */
    uint32_t seed;        /* Seed::instance().set(seed) before the sampler */
    int scheme;           /* VO_* resample scheme */
    double threshold;     /* resample when ESS < threshold * N */
    double dt;            /* 0.1 */
    double sd_pos;        /* process noise sd on positions */
    double sd_vel;        /* process noise sd on velocities */
    double sd_pos0;       /* prior sd of positions (2) */
    double sd_vel0;       /* prior sd of velocities (1) */
    int obs_model;        /* 0 Student-t (reference form, no constants),
                             1 Student-t with normalising constants,
                             2 Gaussian with normalising constants */
    double obs_scale;     /* 0.1 */
    double obs_df;        /* 10 */
    int threads;          /* OpenMP threads for the eval_sp loops (0: all) */
} vo_pf_config;

/* Runs Sampler::initialize + iterate(T-1) (include/vsmc/core/sampler.hpp:
 * 443-476, :735-785) for the constant-velocity PF of
 * user_guide/cpp/pf_seq.cpp with engines from RNGSetVector. Outputs (each
 * may be NULL): ess[T] (ESS before resampling), resampled[T], pos[2T]
 * (monitor "pos" at MonitorMCMC), logz[T] (cumulative log-evidence),
 * state_out[4N] (AoS x,y,vx,vy), W_out[N]. Returns the seconds spent in
 * the timed region. */
/* Per-phase seconds of the last vo_pf_run: move, weights, resampling
 * (counts + rep -> index), copy, monitor. */
void vo_pf_phase_seconds(double out[5]);
double vo_pf_run(const vo_pf_config *cfg, const double *obs_x,
    const double *obs_y, double *ess, int64_t *resampled, double *pos,
    double *logz, double *state_out, double *W_out);

/* FP64 Kalman filter log-likelihood for the linear-Gaussian variant of the
 * same model (obs = position + N(0, sd_obs^2)). Returns log p(y_0:T-1);
 * logz_out[t] (if not NULL) receives the cumulative value. */
double vo_kalman_cv(int64_t T, double dt, double sd_pos, double sd_vel,
    double sd_pos0, double sd_vel0, double sd_obs, const double *obs_x,
    const double *obs_y, double *logz_out);

/* Observation log-density of the PF model at position (x, y) (obs_model
 * as in vo_pf_config). */
double vo_cv_loglik(double x, double y, double ox, double oy, int obs_model,
    double obs_scale, double obs_df);

/* ----------------------------------- tempered mixture SMC (gmm.c) -- */

void vo_gmm_hyper(const double *x, int64_t n, double h[4]);
double vo_gmm_log_prior(int c, const double *mu, const double *lambda,
    const double h[4]);
double vo_gmm_log_lik(int c, const double *mu, const double *lambda,
    const double *weight, const double *x, int64_t n);
void vo_gmm_scales(double alpha, double sd[3]);
int vo_gmm_move(int kind, int c, double *p, const double *z, double log_u,
    double alpha, const double h[4], const double *x, int64_t n);
int64_t vo_gmm_moves(int kind, int c, int64_t n, double *p, const double *z,
    const double *lu, double alpha, const double h[4], const double *x,
    int64_t nd, int threads);
double vo_cess(int64_t n, const double *W, const double *ll, double d);
double vo_cess_solve(int64_t n, const double *W, const double *ll,
    double dmax, double target, double tol);

/* ---------------------------------------- learned proposal (mlp.c) -- */

int vo_mlp_nparams(void);
double vo_bf16(double x);
void vo_cv_features(double x, double y, double vx, double vy, double ox,
    double oy, double dt, double f[6]);
void vo_mlp_forward(const float *params, const double *shift,
    const double *scale, int64_t n, const double *feat, int round, double *out);
double vo_mlp_loss_grad(const float *params, const double *shift,
    const double *scale, int64_t n, const double *feat, const double *target,
    int round, double *grad);
double vo_cv_learned_particle(const double s0[4], double ox, double oy,
    const double o[8], const double z[4], double dt, double sd_pos,
    double sd_vel, int obs_model, double obs_scale, double obs_df,
    double s1[4]);

/* ------------------------------------------------- random walk (rw.c) --
 * RandomWalk / RandomWalkG / NormalProposal / NormalMVProposal
 * (include/vsmc/rng/random_walk.hpp:49-555). */
#define VO_RW_NORMAL 0     /* N(0,1), NormalProposal(1.5), one-step, ltx */
#define VO_RW_GAMMA 1      /* Gamma(2.5), (0, inf), n-step, ltx */
#define VO_RW_BETA 2       /* Beta(2, 3), (0, 1), one-step, ltx null */
#define VO_RW_NEG_GAMMA 3  /* -Gamma(2.5), (-inf, 0), n-step, ltx null */
#define VO_RW_MV3 4        /* 3-D, NormalMVProposal<3> + bounds, n-step */
#define VO_RW_BLOCK 5      /* 3-D, block x[1..3) NormalMVProposal<2>, n-step */
#define VO_RW_FLAT 6       /* flat, NormalProposal(1.3), one-step, ltx null */
#define VO_RW_FLAT_MV 7    /* flat 3-D, NormalMVProposal<3> chol, one-step */
#define VO_RW_G 8          /* RandomWalkG, g = (x, x^2), one-step */
#define VO_RW_BLOCK1 9     /* 3-D, block x[2] NormalProposal (0, inf) */
#define VO_RW_NSCEN 10
#define VO_RW_OUT 8        /* x[3], log gamma, accepted, g[2], words */
/* normal_proposal_q / _qb / _qa / _qab (:378-410) by flag (lower << 1) +
 * upper; returns log q(y, x) / q(x, y) */
double vo_rw_q(unsigned flag, double x, double z, double a, double b,
    double *y);
unsigned vo_rw_flag(double a, double b);
int vo_rw_dim(int scenario);
void vo_rw_init(int scenario, size_t i, double *x);
/* n particles of a scenario, `iters` moves of `steps` updates each;
 * out[i * VO_RW_OUT + k] */
int vo_rw_run(int scenario, uint32_t seed, size_t n, size_t iters,
    size_t steps, double *out);

#ifdef __cplusplus
}
#endif

#endif /* VSMC_ORACLE_H */
