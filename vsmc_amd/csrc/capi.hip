// capi.hip -- the extern "C" entry points declared in include/vsmch.h.
#include "particle.hpp"
#include "mlp.hpp"
#include "genealogy.hpp"
#include "../../include/vsmch.h"
#include "../../include/vsmch/device_rng.hpp"

#include <cmath>
#include <vector>

using namespace vsmch;

namespace {

// The library stream of the calling host thread (vsmch_set_stream).
thread_local hipStream_t g_stream = nullptr;

Workspace &ws_current()
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    Workspace &ws = default_workspace(dev);
    ws.stream = g_stream;
    return ws;
}

bool valid_kind(int k) { return k == VSMCH_PHILOX4X32 || k == VSMCH_THREEFRY4X64; }

CvModelDev cv_dev(const vsmch_cv_model *m)
{
    CvModelDev d{};
    d.dt = m->dt;
    d.sd_pos = m->sd_pos;
    d.sd_vel = m->sd_vel;
    d.sd_pos0 = m->sd_pos0;
    d.sd_vel0 = m->sd_vel0;
    d.obs_scale = m->obs_scale;
    d.inv_scale = 1 / m->obs_scale;
    d.obs_df = m->obs_df;
    d.inv_df = 1 / m->obs_df;
    d.obs_model = m->obs_model;
    const double two_pi = 6.283185307179586476925286766559005768394;
    if (m->obs_model == 1) {
        const double nu = m->obs_df;
        const double k = std::lgamma(0.5 * (nu + 1)) - std::lgamma(0.5 * nu) -
            0.5 * std::log(nu * 3.141592653589793238462643383279502884197) -
            std::log(m->obs_scale);
        d.ll_const = 2 * k;
    } else if (m->obs_model == 2) {
        d.ll_const = -(2 * (std::log(m->obs_scale) + 0.5 * std::log(two_pi)));
    }
    return d;
}

} // namespace

hipStream_t vsmch::library_stream() { return g_stream; }

struct vsmch_particle {
    ParticleSet *p;
    bool borrowed = false;   // owned by another object (vsmch_gmm)
};

struct vsmch_mlp {
    MlpSet *m;
};

struct vsmch_batch {
    BatchSet *b;
};

struct vsmch_gmm {
    GmmSet *s;
    vsmch_particle view;
};

struct vsmch_rindex {
    RIndex *r;
};

namespace {
// The filter loop behind vsmch_cv_run and vsmch_cv_run_learned.
vsmch_status run_filter(vsmch_particle *p, const vsmch_cv_model *m,
    const MlpSet *net, int64_t T, const double *obs_x_host,
    const double *obs_y_host, int scheme, double threshold, double *ess_host,
    int *resampled_host, double *pos_host, double *logz_host);
} // namespace

extern "C" {

const char *vsmch_last_error(void) { return get_error(); }

int vsmch_version(void) { return 10000; }

vsmch_status vsmch_set_stream(void *stream)
{
    g_stream = static_cast<hipStream_t>(stream);
    return VSMCH_OK;
}

vsmch_status vsmch_synchronize(void)
{
    VSMCH_HIP(hipStreamSynchronize(g_stream));
    return VSMCH_OK;
}

vsmch_status vsmch_rng_words(int kind, uint64_t key0, uint64_t word_offset,
    size_t n, void *out)
{
    if (!valid_kind(kind) || (n > 0 && out == nullptr)) {
        set_error("vsmch_rng_words: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return stream_fill(kind, key0, word_offset, static_cast<long long>(n), out,
        0, g_stream);
}

vsmch_status vsmch_u01_fill(int kind, uint64_t key0, uint64_t word_offset,
    size_t n, double *out)
{
    if (!valid_kind(kind) || (n > 0 && out == nullptr)) {
        set_error("vsmch_u01_fill: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return stream_fill(kind, key0, word_offset, static_cast<long long>(n), out,
        1, g_stream);
}

vsmch_status vsmch_normal_fill(int kind, uint64_t key0, uint64_t word_offset,
    size_t n, double mean, double sd, int form, double *out)
{
    if (!valid_kind(kind) || (form != 0 && form != 1) ||
        (n > 0 && out == nullptr)) {
        set_error("vsmch_normal_fill: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return normal_fill(kind, key0, word_offset, static_cast<long long>(n),
        mean, sd, form, out, g_stream);
}

vsmch_status vsmch_resample_trans_u01_rep(int m, int n, const double *weight,
    const double *u01, int *replication)
{
    if (m < 0 || n < 0 || (m > 0 && (!weight || !replication)) ||
        (n > 0 && m > 1 && !u01)) {
        set_error("vsmch_resample_trans_u01_rep: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return trans_u01_rep(ws_current(), m, n, weight, u01, replication);
}

vsmch_status vsmch_resample_trans_u01_index(int m, int n,
    const double *weight, const double *u01, int *index)
{
    // transform.hpp:75-95: index[j] is the particle u_j lands in, i.e. the
    // counts expanded in source order.
    if (m < 0 || n < 0 || (m > 0 && n > 0 && (!weight || !index))) {
        set_error("vsmch_resample_trans_u01_index: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    if (m == 0 || n == 0)
        return VSMCH_OK;
    Workspace &ws = ws_current();
    if (!ws.ensure(m > n ? m : n))
        return VSMCH_ERR_ALLOC;
    int *rep = nullptr;
    VSMCH_HIP(hipMalloc(&rep, sizeof(int) * m));
    int rc = trans_u01_rep(ws, m, n, weight, u01, rep);
    if (rc == VSMCH_OK)
        rc = trans_rep_index(ws, m, n, rep, index, nullptr, 0, 0, 4, false,
            true);
    (void)hipStreamSynchronize(g_stream);
    (void)hipFree(rep);
    return rc;
}

vsmch_status vsmch_resample_trans_rep_index(int m, int n,
    const int *replication, int *index)
{
    if (m < 0 || n < 0 || (m > 0 && n > 0 && (!replication || !index))) {
        set_error("vsmch_resample_trans_rep_index: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return trans_rep_index(ws_current(), m, n, replication, index, nullptr, 0,
        0, 4, false, false);
}

vsmch_status vsmch_resample_trans_index_rep(int m, int n, const int *index,
    int *replication)
{
    if (m < 0 || n < 0 || (m > 0 && n > 0 && (!replication || !index))) {
        set_error("vsmch_resample_trans_index_rep: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return trans_index_rep(ws_current(), m, n, index, replication);
}

vsmch_status vsmch_resample_trans_residual(int m, int n, const double *weight,
    double *resid, int *integ, int *r_host)
{
    if (m < 0 || n < 0 || !r_host || (m > 0 && (!weight || !resid || !integ))) {
        set_error("vsmch_resample_trans_residual: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    long long R = 0;
    const int rc = trans_residual(ws_current(), m, n, weight, resid, integ, &R);
    *r_host = static_cast<int>(R);
    return rc;
}

vsmch_status vsmch_resample(int m, int n, uint32_t key0,
    uint64_t *word_offset_host, const double *weight, int *replication,
    vsmch_resample_scheme scheme)
{
    if (m < 0 || n < 0 || !word_offset_host || scheme < 0 || scheme > 5 ||
        (m > 0 && (!weight || !replication))) {
        set_error("vsmch_resample: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    if (m == 0) {
        // ResampleSystematic still constructs its sequence (one word)
        return VSMCH_OK;
    }
    Workspace &ws = ws_current();
    int rc = resample_scheme(ws, scheme, m, n, key0, *word_offset_host, weight,
        kWeightArray, replication, false);
    if (rc != VSMCH_OK)
        return rc;
    unsigned long long used = 0;
    rc = read_words_used(ws, &used);
    if (rc != VSMCH_OK)
        return rc;
    *word_offset_host += used;
    return VSMCH_OK;
}

#define VSMCH_RESAMPLE_SCHEME(name, scheme)                                    \
    vsmch_status vsmch_resample_##name(int m, int n, uint32_t key0,            \
        uint64_t *word_offset_host, const double *weight, int *replication)    \
    {                                                                          \
        return vsmch_resample(m, n, key0, word_offset_host, weight,            \
            replication, scheme);                                              \
    }
VSMCH_RESAMPLE_SCHEME(multinomial, VSMCH_RESAMPLE_MULTINOMIAL)
VSMCH_RESAMPLE_SCHEME(stratified, VSMCH_RESAMPLE_STRATIFIED)
VSMCH_RESAMPLE_SCHEME(systematic, VSMCH_RESAMPLE_SYSTEMATIC)
VSMCH_RESAMPLE_SCHEME(residual, VSMCH_RESAMPLE_RESIDUAL)
VSMCH_RESAMPLE_SCHEME(residual_stratified, VSMCH_RESAMPLE_RESIDUAL_STRATIFIED)
VSMCH_RESAMPLE_SCHEME(residual_systematic, VSMCH_RESAMPLE_RESIDUAL_SYSTEMATIC)
#undef VSMCH_RESAMPLE_SCHEME

vsmch_status vsmch_u01_sorted(int n, const double *u01, double *u01seq)
{
    if (n < 0 || (n > 0 && (!u01 || !u01seq))) {
        set_error("vsmch_u01_sorted: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return u01_sorted_seq(ws_current(), n, u01, u01seq);
}

vsmch_status vsmch_u01_stratified(int n, const double *u01, double *u01seq)
{
    if (n < 0 || (n > 0 && (!u01 || !u01seq))) {
        set_error("vsmch_u01_stratified: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return u01_stratified_seq(ws_current(), n, u01, u01seq);
}

vsmch_status vsmch_u01_systematic(int n, double u01, double *u01seq)
{
    if (n < 0 || (n > 0 && !u01seq)) {
        set_error("vsmch_u01_systematic: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return u01_systematic_seq(ws_current(), n, u01, u01seq);
}

vsmch_status vsmch_exact_cumsum(size_t n, const double *w, double *out)
{
    if (n > 0 && (!w || !out)) {
        set_error("vsmch_exact_cumsum: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return exact_cumsum(ws_current(), static_cast<long long>(n), w, out);
}

vsmch_status vsmch_sorted_uniforms(size_t n, uint32_t key0,
    uint64_t word_offset, double *out)
{
    if (n > 0 && !out) {
        set_error("vsmch_sorted_uniforms: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return sorted_uniforms(ws_current(), static_cast<long long>(n), key0,
        word_offset, out);
}

// ------------------------------------------------------------ particles

vsmch_status vsmch_particle_create(vsmch_particle **out,
    const vsmch_particle_config *cfg)
{
    if (!out || !cfg) {
        set_error("vsmch_particle_create: null argument");
        return VSMCH_ERR_INVALID;
    }
    ParticleSet *p = nullptr;
    const int rc = particle_create(&p, cfg->n, cfg->dim, cfg->state_bytes,
        cfg->seed, g_stream);
    if (rc != VSMCH_OK)
        return rc;
    *out = new vsmch_particle{p};
    return VSMCH_OK;
}

vsmch_status vsmch_particle_destroy(vsmch_particle *p)
{
    if (!p)
        return VSMCH_OK;
    if (p->borrowed) {
        set_error("vsmch_particle_destroy: a borrowed particle set belongs to its sampler");
        return VSMCH_ERR_INVALID;
    }
    (void)hipStreamSynchronize(p->p->ws.stream);
    particle_destroy(p->p);
    delete p;
    return VSMCH_OK;
}

vsmch_status vsmch_particle_state(vsmch_particle *p, void **state)
{
    if (!p || !state)
        return VSMCH_ERR_INVALID;
    *state = p->p->state;
    return VSMCH_OK;
}

vsmch_status vsmch_particle_log_weight(vsmch_particle *p, double **logw)
{
    if (!p || !logw)
        return VSMCH_ERR_INVALID;
    *logw = p->p->logw;
    return VSMCH_OK;
}

vsmch_status vsmch_particle_read_state(vsmch_particle *p, void *dst)
{
    if (!p || !dst)
        return VSMCH_ERR_INVALID;
    ParticleSet *ps = p->p;
    VSMCH_HIP(hipMemcpyAsync(dst, ps->state,
        static_cast<size_t>(ps->n) * ps->dim * ps->bytes, hipMemcpyDeviceToDevice,
        ps->ws.stream));
    VSMCH_HIP(hipStreamSynchronize(ps->ws.stream));
    return VSMCH_OK;
}

vsmch_status vsmch_particle_write_state(vsmch_particle *p, const void *src)
{
    if (!p || !src)
        return VSMCH_ERR_INVALID;
    ParticleSet *ps = p->p;
    VSMCH_HIP(hipMemcpyAsync(ps->state, src,
        static_cast<size_t>(ps->n) * ps->dim * ps->bytes, hipMemcpyDeviceToDevice,
        ps->ws.stream));
    VSMCH_HIP(hipStreamSynchronize(ps->ws.stream));
    return VSMCH_OK;
}

vsmch_status vsmch_particle_keys(vsmch_particle *p, uint32_t *key_first_host,
    uint32_t *resample_key_host, uint64_t *resample_word_host)
{
    if (!p)
        return VSMCH_ERR_INVALID;
    PDev h{};
    const int rc = read_pdev(p->p, &h);
    if (rc != VSMCH_OK)
        return rc;
    if (key_first_host)
        *key_first_host = seed_key(p->p->seed, 0);
    if (resample_key_host)
        *resample_key_host = h.rkey;
    if (resample_word_host)
        *resample_word_host = h.rword;
    return VSMCH_OK;
}

vsmch_status vsmch_weight_set_equal(vsmch_particle *p)
{
    if (!p)
        return VSMCH_ERR_INVALID;
    return weight_update(p->p, nullptr, 2);
}

vsmch_status vsmch_weight_set_log(vsmch_particle *p, const double *logw)
{
    if (!p || !logw)
        return VSMCH_ERR_INVALID;
    return weight_update(p->p, logw, 0);
}

vsmch_status vsmch_weight_add_log(vsmch_particle *p, const double *inc)
{
    if (!p || !inc)
        return VSMCH_ERR_INVALID;
    return weight_update(p->p, inc, 1);
}

vsmch_status vsmch_weight_update(vsmch_particle *p, int op, const double *src,
    int64_t stride)
{
    if (!p || !src || stride < 1 ||
        (op != VSMCH_WEIGHT_SET_LOG && op != VSMCH_WEIGHT_ADD_LOG &&
         op != VSMCH_WEIGHT_SET && op != VSMCH_WEIGHT_MUL)) {
        set_error("vsmch_weight_update: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return weight_update(p->p, src, op, stride);
}

vsmch_status vsmch_particle_copy(vsmch_particle *p, int64_t n, const int *index)
{
    if (!p || !index || n != p->p->n) {
        set_error("vsmch_particle_copy: invalid argument (n must equal N)");
        return VSMCH_ERR_INVALID;
    }
    return state_copy(p->p, index);
}

vsmch_status vsmch_weight_ess(vsmch_particle *p, double *ess_host)
{
    if (!p || !ess_host)
        return VSMCH_ERR_INVALID;
    PDev h{};
    const int rc = read_pdev(p->p, &h);
    *ess_host = h.ess;
    return rc;
}

vsmch_status vsmch_weight_log_zconst(vsmch_particle *p, double *logz_host)
{
    if (!p || !logz_host)
        return VSMCH_ERR_INVALID;
    PDev h{};
    const int rc = read_pdev(p->p, &h);
    *logz_host = h.logz;
    return rc;
}

vsmch_status vsmch_weight_read(vsmch_particle *p, double *W)
{
    if (!p || !W)
        return VSMCH_ERR_INVALID;
    return weight_read(p->p, W);
}

vsmch_status vsmch_particle_resample(vsmch_particle *p, int scheme,
    double threshold, int *resampled_host, int *rep_out, int *idx_out)
{
    if (!p || scheme < 0 || scheme > 5)
        return VSMCH_ERR_INVALID;
    ParticleSet *ps = p->p;
    // with a host result the decision is read first: a step that does not
    // resample then queues no further work
    int rc = resample_step(ps, scheme, threshold, 0, nullptr, nullptr, rep_out,
        idx_out, resampled_host != nullptr);
    if (rc != VSMCH_OK)
        return rc;
    if (resampled_host) {
        PDev h{};
        rc = read_pdev(ps, &h);
        *resampled_host = h.resampled;
    }
    return rc;
}

vsmch_status vsmch_particle_rng_words(vsmch_particle *p, uint64_t **words)
{
    if (!p || !words)
        return VSMCH_ERR_INVALID;
    ParticleSet *ps = p->p;
    if (!ps->rng_words) {
        VSMCH_HIP(hipMalloc(&ps->rng_words, sizeof(unsigned long long) * ps->n));
        VSMCH_HIP(hipMemset(ps->rng_words, 0, sizeof(unsigned long long) * ps->n));
    }
    *words = reinterpret_cast<uint64_t *>(ps->rng_words);
    return VSMCH_OK;
}

vsmch_status vsmch_particle_scratch(vsmch_particle *p, double **scratch)
{
    if (!p || !scratch)
        return VSMCH_ERR_INVALID;
    *scratch = p->p->scratch;
    return VSMCH_OK;
}

vsmch_status vsmch_weighted_sum(vsmch_particle *p, int dim, const double *buf,
    double *r_host)
{
    if (!p || !buf || !r_host || dim < 1)
        return VSMCH_ERR_INVALID;
    ParticleSet *ps = p->p;
    double *out = nullptr;
    VSMCH_HIP(hipMalloc(&out, sizeof(double) * dim));
    int rc = weighted_sum(ps, dim, buf, out);
    if (rc == VSMCH_OK) {
        VSMCH_HIP(hipMemcpyAsync(r_host, out, sizeof(double) * dim,
            hipMemcpyDeviceToHost, ps->ws.stream));
        VSMCH_HIP(hipStreamSynchronize(ps->ws.stream));
    }
    (void)hipFree(out);
    return rc;
}

vsmch_status vsmch_particle_resample_count(vsmch_particle *p,
    int64_t *count_host)
{
    if (!p || !count_host)
        return VSMCH_ERR_INVALID;
    PDev h{};
    const int rc = read_pdev(p->p, &h);
    *count_host = h.nresampled;
    return rc;
}

vsmch_status vsmch_monitor_state_moments(vsmch_particle *p, int c0, int dim,
    double *r_host)
{
    if (!p || !r_host)
        return VSMCH_ERR_INVALID;
    int rc = moments(p->p, c0, dim, p->p->mout);
    if (rc != VSMCH_OK)
        return rc;
    VSMCH_HIP(hipMemcpyAsync(r_host, p->p->mout, sizeof(double) * dim,
        hipMemcpyDeviceToHost, p->p->ws.stream));
    VSMCH_HIP(hipStreamSynchronize(p->p->ws.stream));
    return VSMCH_OK;
}

vsmch_status vsmch_particle_cess(vsmch_particle *p, const double *ll,
    double d, double *cess_host)
{
    if (!p || !ll || !cess_host)
        return VSMCH_ERR_INVALID;
    return particle_cess(p->p, ll, d, cess_host);
}

vsmch_status vsmch_particle_cess_solve(vsmch_particle *p, const double *ll,
    double target, double dmax, int rounds, double *d_host)
{
    if (!p || !ll || !d_host)
        return VSMCH_ERR_INVALID;
    return particle_cess_solve(p->p, ll, target, dmax, rounds, d_host);
}

vsmch_status vsmch_cv_init(vsmch_particle *p, const vsmch_cv_model *m,
    double obs_x, double obs_y)
{
    if (!p || !m)
        return VSMCH_ERR_INVALID;
    return cv_step(p->p, cv_dev(m), 0, obs_x, obs_y, false);
}

vsmch_status vsmch_cv_move(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t iter, double obs_x, double obs_y)
{
    if (!p || !m || iter < 1)
        return VSMCH_ERR_INVALID;
    return cv_step(p->p, cv_dev(m), iter, obs_x, obs_y, false);
}

vsmch_status vsmch_cv_run(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t T, const double *obs_x_host, const double *obs_y_host,
    int scheme, double threshold, double *ess_host, int *resampled_host,
    double *pos_host, double *logz_host)
{
    return run_filter(p, m, nullptr, T, obs_x_host, obs_y_host, scheme,
        threshold, ess_host, resampled_host, pos_host, logz_host);
}

vsmch_status vsmch_cv_iterate(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t iter0, int64_t k, const double *obs_x_host,
    const double *obs_y_host, int scheme, double threshold, double *ess,
    int *resampled, double *pos, double *logz)
{
    if (!p || !m || iter0 < 0 || k < 0 || scheme < 0 || scheme > 5 ||
        (k > 0 && (!obs_x_host || !obs_y_host))) {
        set_error("vsmch_cv_iterate: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return cv_iterate(p->p, cv_dev(m), iter0, k, obs_x_host, obs_y_host,
        scheme, threshold, ess, resampled, pos, logz);
}

vsmch_status vsmch_mlp_create(vsmch_mlp **m)
{
    if (!m)
        return VSMCH_ERR_INVALID;
    MlpSet *s = nullptr;
    const int rc = mlp_create(&s, g_stream);
    if (rc != VSMCH_OK)
        return rc;
    *m = new vsmch_mlp{s};
    return VSMCH_OK;
}

vsmch_status vsmch_mlp_destroy(vsmch_mlp *m)
{
    if (!m)
        return VSMCH_ERR_INVALID;
    mlp_destroy(m->m);
    delete m;
    return VSMCH_OK;
}

vsmch_status vsmch_mlp_set_params(vsmch_mlp *m, const float *params_host,
    const float *shift_host, const float *scale_host)
{
    if (!m)
        return VSMCH_ERR_INVALID;
    return mlp_set_params(m->m, params_host, shift_host, scale_host);
}

vsmch_status vsmch_mlp_get_params(vsmch_mlp *m, float *params_host)
{
    if (!m || !params_host)
        return VSMCH_ERR_INVALID;
    return mlp_get_params(m->m, params_host);
}

vsmch_status vsmch_mlp_forward(vsmch_mlp *m, int64_t n, const float *feat,
    float *out)
{
    if (!m || n < 0 || (n > 0 && (!feat || !out)))
        return VSMCH_ERR_INVALID;
    return mlp_forward(m->m, n, feat, out);
}

vsmch_status vsmch_mlp_loss_grad(vsmch_mlp *m, int64_t n, const float *feat,
    const float *target, float *grad, float *loss_host)
{
    if (!m || n < 1 || !feat || !target || !grad)
        return VSMCH_ERR_INVALID;
    return mlp_loss_grad(m->m, n, feat, target, grad, loss_host);
}

vsmch_status vsmch_mlp_train(vsmch_mlp *m, int64_t n, const float *feat,
    const float *target, int64_t batch, int64_t steps, float lr,
    float *loss_host)
{
    if (!m || !feat || !target)
        return VSMCH_ERR_INVALID;
    return mlp_train(m->m, n, feat, target, batch, steps, lr, loss_host);
}

vsmch_status vsmch_cv_transitions(vsmch_particle *p, const vsmch_cv_model *m,
    int64_t T, const double *obs_x_host, const double *obs_y_host,
    int scheme, double threshold, uint32_t train_seed, float *feat,
    float *target)
{
    if (!p || !m || T < 2 || !obs_x_host || !obs_y_host || scheme < 0 ||
        scheme > 5 || !feat || !target)
        return VSMCH_ERR_INVALID;
    const int rc = cv_transitions(p->p, cv_dev(m), T, obs_x_host, obs_y_host,
        scheme, threshold, train_seed, feat, target);
    if (rc != VSMCH_OK)
        return rc;
    VSMCH_HIP(hipStreamSynchronize(p->p->ws.stream));
    return VSMCH_OK;
}

vsmch_status vsmch_cv_move_learned(vsmch_particle *p, const vsmch_cv_model *m,
    const vsmch_mlp *net, int64_t iter, double obs_x, double obs_y)
{
    if (!p || !m || !net || iter < 1)
        return VSMCH_ERR_INVALID;
    return cv_step_learned(p->p, cv_dev(m), net->m, iter, obs_x, obs_y);
}

vsmch_status vsmch_batch_create(vsmch_batch **b, int nfilters,
    const vsmch_batch_filter *filters_host, const vsmch_cv_model *m,
    int scheme, double threshold)
{
    if (!b || !filters_host || !m || nfilters < 1) {
        set_error("vsmch_batch_create: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    std::vector<long long> n(nfilters), T(nfilters);
    std::vector<uint32_t> seed(nfilters);
    for (int f = 0; f < nfilters; ++f) {
        n[f] = filters_host[f].n;
        T[f] = filters_host[f].T;
        seed[f] = filters_host[f].seed;
    }
    BatchSet *s = nullptr;
    const int rc = batch_create(&s, nfilters, n.data(), T.data(), seed.data(),
        cv_dev(m), scheme, threshold);
    if (rc != VSMCH_OK)
        return rc;
    *b = new vsmch_batch{s};
    return VSMCH_OK;
}

vsmch_status vsmch_batch_destroy(vsmch_batch *b)
{
    if (!b)
        return VSMCH_ERR_INVALID;
    batch_destroy(b->b);
    delete b;
    return VSMCH_OK;
}

vsmch_status vsmch_batch_set_obs(vsmch_batch *b, const double *obs_host)
{
    if (!b || !obs_host)
        return VSMCH_ERR_INVALID;
    return batch_set_obs(b->b, obs_host);
}

vsmch_status vsmch_batch_run(vsmch_batch *b, const vsmch_mlp *net,
    int use_graph)
{
    if (!b)
        return VSMCH_ERR_INVALID;
    return batch_run(b->b, net ? net->m : nullptr, use_graph != 0);
}

vsmch_status vsmch_batch_histories(vsmch_batch *b, double *logz_host,
    double *ess_host, int *resampled_host)
{
    if (!b)
        return VSMCH_ERR_INVALID;
    return batch_histories(b->b, logz_host, ess_host, resampled_host);
}

vsmch_status vsmch_batch_read_filter(vsmch_batch *b, int f, float *state,
    double *weight)
{
    if (!b)
        return VSMCH_ERR_INVALID;
    return batch_read_filter(b->b, f, state, weight);
}

vsmch_status vsmch_batch_info(vsmch_batch *b, int64_t *tmax,
    int64_t *particles, int *move_blocks, double *run_ms)
{
    if (!b)
        return VSMCH_ERR_INVALID;
    long long tm = 0, np = 0;
    batch_info(b->b, &tm, &np, move_blocks, run_ms);
    if (tmax)
        *tmax = tm;
    if (particles)
        *particles = np;
    return VSMCH_OK;
}

vsmch_status vsmch_gmm_create(vsmch_gmm **g, const vsmch_gmm_config *cfg,
    const double *data_host, int64_t ndata)
{
    if (!g || !cfg || !data_host) {
        set_error("vsmch_gmm_create: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    GmmSet *s = nullptr;
    const int rc = gmm_create(&s, cfg->n, cfg->ncomp, cfg->seed, cfg->cess,
        cfg->moves, cfg->search_rounds, cfg->scheme, cfg->threshold, data_host,
        ndata, g_stream);
    if (rc != VSMCH_OK)
        return rc;
    *g = new vsmch_gmm{s, vsmch_particle{gmm_particles(s), true}};
    return VSMCH_OK;
}

vsmch_status vsmch_gmm_destroy(vsmch_gmm *g)
{
    if (!g)
        return VSMCH_ERR_INVALID;
    gmm_destroy(g->s);
    delete g;
    return VSMCH_OK;
}

vsmch_status vsmch_gmm_run(vsmch_gmm *g, int64_t max_steps,
    int64_t *steps_host, double *alpha_host, double *ess_host,
    int *resampled_host, double *logz_host, double *integrand_host,
    double *accept_host)
{
    if (!g || max_steps < 1) {
        set_error("vsmch_gmm_run: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    const size_t T = static_cast<size_t>(max_steps);
    double *a = nullptr, *e = nullptr, *z = nullptr, *in = nullptr, *acc = nullptr;
    int *rs = nullptr;
    if (hipMalloc(&a, sizeof(double) * T) != hipSuccess ||
        hipMalloc(&e, sizeof(double) * T) != hipSuccess ||
        hipMalloc(&z, sizeof(double) * T) != hipSuccess ||
        hipMalloc(&in, sizeof(double) * T) != hipSuccess ||
        hipMalloc(&acc, sizeof(double) * 3 * T) != hipSuccess ||
        hipMalloc(&rs, sizeof(int) * T) != hipSuccess) {
        void *ps[] = {a, e, z, in, acc, rs};
        for (void *p : ps)
            (void)hipFree(p);
        set_error("vsmch_gmm_run: device allocation failed");
        return VSMCH_ERR_ALLOC;
    }
    long long steps = 0;
    int rc = gmm_run(g->s, max_steps, a, e, rs, z, in, acc, &steps);
    if (rc == VSMCH_OK) {
        const size_t k = static_cast<size_t>(steps);
        struct {
            void *dst;
            const void *src;
            size_t bytes;
        } cp[] = {{alpha_host, a, sizeof(double) * k}, {ess_host, e, sizeof(double) * k},
            {resampled_host, rs, sizeof(int) * k}, {logz_host, z, sizeof(double) * k},
            {integrand_host, in, sizeof(double) * k},
            {accept_host, acc, sizeof(double) * 3 * k}};
        for (auto &c : cp)
            if (c.dst && c.bytes &&
                hipMemcpy(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost) != hipSuccess) {
                set_error("vsmch_gmm_run: copy failed");
                rc = VSMCH_ERR_HIP;
            }
        if (steps_host)
            *steps_host = steps;
    }
    void *ps[] = {a, e, z, in, acc, rs};
    for (void *p : ps)
        (void)hipFree(p);
    return rc;
}

vsmch_status vsmch_gmm_particle(vsmch_gmm *g, vsmch_particle **p)
{
    if (!g || !p)
        return VSMCH_ERR_INVALID;
    *p = &g->view;
    return VSMCH_OK;
}

vsmch_status vsmch_gmm_move(vsmch_gmm *g, int kind, double alpha,
    int64_t iter, int m)
{
    if (!g || kind < 0 || kind > 2 || iter < 0 || m < 0 || m >= 32)
        return VSMCH_ERR_INVALID;
    int rc = gmm_set_alpha(g->s, alpha);
    if (rc != VSMCH_OK)
        return rc;
    return gmm_move(g->s, kind, iter, m);
}

vsmch_status vsmch_cv_run_learned(vsmch_particle *p, const vsmch_cv_model *m,
    const vsmch_mlp *net, int64_t T, const double *obs_x_host,
    const double *obs_y_host, int scheme, double threshold, double *ess_host,
    int *resampled_host, double *pos_host, double *logz_host)
{
    if (!net)
        return VSMCH_ERR_INVALID;
    return run_filter(p, m, net->m, T, obs_x_host, obs_y_host, scheme,
        threshold, ess_host, resampled_host, pos_host, logz_host);
}

// ---------------------------------------------------------- genealogy

vsmch_status vsmch_rindex_create(vsmch_rindex **r, int64_t n)
{
    if (!r)
        return VSMCH_ERR_INVALID;
    RIndex *x = nullptr;
    const int rc = rindex_create(&x, n, g_stream);
    if (rc != VSMCH_OK)
        return rc;
    *r = new vsmch_rindex{x};
    return VSMCH_OK;
}

vsmch_status vsmch_rindex_destroy(vsmch_rindex *r)
{
    if (!r)
        return VSMCH_ERR_INVALID;
    rindex_destroy(r->r);
    delete r;
    return VSMCH_OK;
}

vsmch_status vsmch_rindex_iter_size(vsmch_rindex *r, int64_t *iters)
{
    if (!r || !iters)
        return VSMCH_ERR_INVALID;
    *iters = r->r->iters;
    return VSMCH_OK;
}

vsmch_status vsmch_rindex_reset(vsmch_rindex *r)
{
    if (!r)
        return VSMCH_ERR_INVALID;
    r->r->iters = 0;
    return VSMCH_OK;
}

vsmch_status vsmch_rindex_clear(vsmch_rindex *r)
{
    if (!r)
        return VSMCH_ERR_INVALID;
    return rindex_clear(r->r);
}

vsmch_status vsmch_rindex_push_back(vsmch_rindex *r, const int *idx)
{
    if (!r)
        return VSMCH_ERR_INVALID;
    return rindex_push_back(r->r, idx);
}

vsmch_status vsmch_rindex_insert(vsmch_rindex *r, int64_t iter, const int *idx)
{
    if (!r)
        return VSMCH_ERR_INVALID;
    return rindex_insert(r->r, iter, idx);
}

vsmch_status vsmch_rindex_index(vsmch_rindex *r, int64_t id, int64_t iter,
    int64_t *idx_host)
{
    if (!r || !idx_host)
        return VSMCH_ERR_INVALID;
    long long v = 0;
    const int rc = rindex_index(r->r, id, iter, &v);
    if (rc == VSMCH_OK)
        *idx_host = v;
    return rc;
}

vsmch_status vsmch_rindex_index_matrix(vsmch_rindex *r, int layout, int *out)
{
    if (!r || !out || (layout != VSMCH_ROW_MAJOR && layout != VSMCH_COL_MAJOR)) {
        set_error("vsmch_rindex_index_matrix: invalid argument");
        return VSMCH_ERR_INVALID;
    }
    return rindex_matrix(r->r, layout == VSMCH_COL_MAJOR ? 1 : 0, out);
}

} // extern "C"

namespace {

vsmch_status run_filter(vsmch_particle *p, const vsmch_cv_model *m,
    const MlpSet *net, int64_t T, const double *obs_x_host,
    const double *obs_y_host, int scheme, double threshold, double *ess_host,
    int *resampled_host, double *pos_host, double *logz_host)
{
    if (!p || !m || T < 1 || !obs_x_host || !obs_y_host || scheme < 0 ||
        scheme > 5)
        return VSMCH_ERR_INVALID;
    ParticleSet *ps = p->p;
    hipStream_t st = ps->ws.stream;
    double *ess_d = nullptr, *pos_d = nullptr, *logz_d = nullptr;
    int *rs_d = nullptr;
    if (hipMalloc(&ess_d, sizeof(double) * T) != hipSuccess ||
        hipMalloc(&pos_d, sizeof(double) * 2 * T) != hipSuccess ||
        hipMalloc(&logz_d, sizeof(double) * T) != hipSuccess ||
        hipMalloc(&rs_d, sizeof(int) * T) != hipSuccess) {
        (void)hipFree(ess_d);
        (void)hipFree(pos_d);
        (void)hipFree(logz_d);
        (void)hipFree(rs_d);
        set_error("filter run: device allocation failed");
        return VSMCH_ERR_ALLOC;
    }
    int rc = cv_run(ps, cv_dev(m), T, obs_x_host, obs_y_host, scheme,
        threshold, ess_d, rs_d, pos_d, logz_d, net);
    if (rc == VSMCH_OK) {
        if (ess_host)
            (void)hipMemcpyAsync(ess_host, ess_d, sizeof(double) * T,
                hipMemcpyDeviceToHost, st);
        if (resampled_host)
            (void)hipMemcpyAsync(resampled_host, rs_d, sizeof(int) * T,
                hipMemcpyDeviceToHost, st);
        if (pos_host)
            (void)hipMemcpyAsync(pos_host, pos_d, sizeof(double) * 2 * T,
                hipMemcpyDeviceToHost, st);
        if (logz_host)
            (void)hipMemcpyAsync(logz_host, logz_d, sizeof(double) * T,
                hipMemcpyDeviceToHost, st);
        if (hipStreamSynchronize(st) != hipSuccess) {
            set_error("vsmch_cv_run: stream error");
            rc = VSMCH_ERR_HIP;
        }
    }
    (void)hipFree(ess_d);
    (void)hipFree(pos_d);
    (void)hipFree(logz_d);
    (void)hipFree(rs_d);
    return rc;
}

} // namespace
