/* rw.c -- CPU oracle (test infrastructure only) of vSMC's random-walk
 * Metropolis-Hastings library, include/vsmc/rng/random_walk.hpp:
 *   RandomWalk      (:49-191)  one-step, block, multi-step updates
 *   RandomWalkG     (:192-356) the same with a test function g(x)
 *   NormalProposal  (:414-452) and its bounded transforms (:378-410)
 *   NormalMVProposal(:456-555) with NormalMVDistribution's generate
 *                   (include/vsmc/rng/normal_mv_distribution.hpp:336-343:
 *                   dim scalar cached normals, normal_distribution.hpp:
 *                   127-136, then the packed lower Cholesky product)
 * restated over a per-particle Philox engine (vo_philox), counting the words
 * each particle consumes.
 *
 * vo_rw_run runs one of the fixed scenarios below for n particles; the
 * device test program (tests/cpp/rw_hip.cpp) runs the same scenarios through
 * the Sampler layer with include/vsmch/random_walk.hpp.
 *
 * Parity: pinned by (a) a flat target, where every proposal is accepted and
 * the chain is x0 plus the stream's Box-Muller draws in order (checked
 * against normals built directly from vo_philox_words), (b) the transforms'
 * closed forms (logit / log shift by z, the log q-ratio as the Jacobian),
 * (c) the stationary moments of long chains against the targets' known
 * means and variances, (d) the Cholesky product on a flat target against
 * numpy. The BLAS tpmv summation order of the reference is not pinned:
 * both sides here sum each row left to right. The reference's own
 * random_walk.hpp cannot be compiled here (its include chain reaches
 * include/vsmc/math/cblas.h, which needs a system <cblas.h> or MKL). */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* ------------------------------------------------ engine with a counter */
typedef struct {
    vo_philox e;
    uint64_t used;
} rw_eng;

static uint32_t rw_next(rw_eng *g)
{
    g->used += 1;
    return vo_philox_next(&g->e);
}

/* NormalDistribution scalar generate with the cached half; counts the two
 * words of a fresh pair */
static double rw_normal(vo_normal *d, rw_eng *g)
{
    if (!d->saved)
        g->used += 2;
    return vo_normal_philox(d, &g->e);
}

static double rw_u01(rw_eng *g) { return vo_u01_co_u32d(rw_next(g)); }

/* ------------------------------------------- transforms (:378-410) */
double vo_rw_q(unsigned flag, double x, double z, double a, double b,
    double *y)
{
    switch (flag) {
    case 0: /* normal_proposal_q */
        *y = x + z;
        return 0;
    case 1: /* normal_proposal_qb */
        *y = b - (b - x) * exp(z);
        return z;
    case 2: /* normal_proposal_qa */
        *y = a + (x - a) * exp(z);
        return z;
    case 3: { /* normal_proposal_qab */
        const double r = exp(z) * (x - a) / (b - x);
        *y = (a + b * r) / (1 + r);
        return log((*y - a) / (x - a) * (b - *y) / (b - x));
    }
    default: return 0;
    }
}

unsigned vo_rw_flag(double a, double b)
{
    const unsigned lower = isfinite(a) ? 1u : 0u;
    const unsigned upper = isfinite(b) ? 1u : 0u;
    return (lower << 1) + upper;
}

/* ----------------------------------------------------- proposals */
#define RW_MAXD 4

typedef struct {
    int mv;        /* 0: NormalProposal, 1: NormalMVProposal */
    size_t dim;
    vo_normal rnorm;
    double chol[RW_MAXD * (RW_MAXD + 1) / 2];
    int has_chol;
    double a[RW_MAXD], b[RW_MAXD];
    unsigned flag[RW_MAXD];
} rw_prop;

static void prop_normal(rw_prop *p, double sd, double a, double b)
{
    memset(p, 0, sizeof(*p));
    p->dim = 1;
    vo_normal_init(&p->rnorm, 0, sd);
    p->a[0] = a;
    p->b[0] = b;
    p->flag[0] = vo_rw_flag(a, b);
}

static void prop_mv(rw_prop *p, size_t dim, const double *chol,
    const double *a, const double *b)
{
    memset(p, 0, sizeof(*p));
    p->mv = 1;
    p->dim = dim;
    vo_normal_init(&p->rnorm, 0, 1);
    p->has_chol = chol != NULL;
    if (chol)
        memcpy(p->chol, chol, sizeof(double) * dim * (dim + 1) / 2);
    for (size_t i = 0; i < dim; ++i) {
        p->a[i] = a ? a[i] : -INFINITY;
        p->b[i] = b ? b[i] : INFINITY;
        p->flag[i] = vo_rw_flag(p->a[i], p->b[i]);
    }
}

/* proposal(rng, dim, x, y): returns log q(y, x) / q(x, y) */
static double prop_apply(rw_prop *p, rw_eng *g, const double *x, double *y)
{
    if (!p->mv)
        return vo_rw_q(p->flag[0], x[0], rw_normal(&p->rnorm, g), p->a[0],
            p->b[0], y);
    double z[RW_MAXD];
    for (size_t i = 0; i < p->dim; ++i)
        z[i] = rw_normal(&p->rnorm, g);
    if (p->has_chol) {
        /* z = L z, L lower triangular packed by rows */
        for (size_t i = p->dim; i-- > 0;) {
            double s = 0;
            for (size_t j = 0; j <= i; ++j)
                s += p->chol[i * (i + 1) / 2 + j] * z[j];
            z[i] = s;
        }
    }
    double q = 0;
    for (size_t i = 0; i < p->dim; ++i)
        q += vo_rw_q(p->flag[i], x[i], z[i], p->a[i], p->b[i], &y[i]);
    return q;
}

/* ------------------------------------------------------- targets */
typedef double (*rw_target)(size_t m, const double *x, double *g);

static double lt_normal(size_t m, const double *x, double *g)
{
    (void)m;
    if (g) {
        g[0] = x[0];
        g[1] = x[0] * x[0];
    }
    return -0.5 * x[0] * x[0];
}

/* Gamma(2.5, 1) on (0, inf) */
static double lt_gamma(size_t m, const double *x, double *g)
{
    (void)m;
    (void)g;
    return 1.5 * log(x[0]) - x[0];
}

/* Beta(2, 3) on (0, 1) */
static double lt_beta(size_t m, const double *x, double *g)
{
    (void)m;
    (void)g;
    return log(x[0]) + 2 * log(1 - x[0]);
}

/* -Gamma(2.5, 1) on (-inf, 0) */
static double lt_neg_gamma(size_t m, const double *x, double *g)
{
    (void)m;
    (void)g;
    return 1.5 * log(-x[0]) + x[0];
}

/* x0 ~ N(0, 1), x1 | x0 ~ N(x0 / 2, 1/2), x2 ~ Gamma(2, 1) */
static double lt_mv3(size_t m, const double *x, double *g)
{
    (void)m;
    (void)g;
    const double d = x[1] - 0.5 * x[0];
    const double l2 = x[2] > 0 ? log(x[2]) - x[2] : -INFINITY;
    return -0.5 * (x[0] * x[0] + d * d / 0.5) + l2;
}

static double lt_flat(size_t m, const double *x, double *g)
{
    (void)m;
    (void)x;
    (void)g;
    return 0;
}

/* ------------------------------------- RandomWalk(G) (:97-188, :260-350) */
typedef struct {
    size_t dim;    /* dimension of the proposal (block) */
    size_t dim_g;  /* 0: RandomWalk, else RandomWalkG */
} rw_walk;

/* one-step update of x[idx .. idx + dim) of a state of length m */
static int rw_step(const rw_walk *w, rw_eng *g, size_t m, size_t idx,
    double *x, double *ltx, double *gout, rw_target lt, rw_prop *p)
{
    double xo[RW_MAXD], y[RW_MAXD], gy[RW_MAXD];
    double xs[RW_MAXD];
    memcpy(xo, x + idx, sizeof(double) * w->dim);
    const double q = prop_apply(p, g, xo, y);
    const double s = ltx == NULL ? lt(m, x, NULL) : *ltx;
    memcpy(xs, x, sizeof(double) * m);
    memcpy(xs + idx, y, sizeof(double) * w->dim);
    const double t = lt(m, xs, gout == NULL ? NULL : gy);
    const double pr = t - s + q;
    const double u = log(rw_u01(g));
    if (u < pr) {
        memcpy(x + idx, y, sizeof(double) * w->dim);
        if (ltx != NULL)
            *ltx = t;
        if (gout != NULL)
            memcpy(gout, gy, sizeof(double) * w->dim_g);
        return 1;
    }
    return 0;
}

/* n-step update (:151-188): log gamma(x) carried between steps */
static int rw_steps(const rw_walk *w, size_t n, rw_eng *g, size_t m,
    size_t idx, double *x, double *ltx, double *gout, rw_target lt,
    rw_prop *p)
{
    int acc = 0;
    double s = ltx == NULL ? lt(m, x, NULL) : *ltx;
    for (size_t i = 0; i < n; ++i)
        acc += rw_step(w, g, m, idx, x, &s, gout, lt, p);
    if (ltx != NULL)
        *ltx = s;
    return acc;
}

/* ------------------------------------------------------- scenarios */
static const double kChol3[6] = {0.5, 0.2, 0.4, 0.1, -0.1, 0.3};
static const double kA3[3] = {-INFINITY, -6, 0};
static const double kB3[3] = {INFINITY, 6, INFINITY};
static const double kChol2[3] = {0.4, 0.1, 0.3};
static const double kA2[2] = {-6, 0};
static const double kB2[2] = {6, INFINITY};

int vo_rw_dim(int scenario)
{
    switch (scenario) {
    case VO_RW_MV3: case VO_RW_BLOCK: case VO_RW_FLAT_MV: case VO_RW_BLOCK1:
        return 3;
    case VO_RW_NORMAL: case VO_RW_GAMMA: case VO_RW_BETA: case VO_RW_NEG_GAMMA:
    case VO_RW_FLAT: case VO_RW_G:
        return 1;
    default: return 0;
    }
}

/* initial state of particle i (no random draws) */
void vo_rw_init(int scenario, size_t i, double *x)
{
    switch (scenario) {
    case VO_RW_NORMAL: x[0] = -2 + 4.0 * (double)(i % 97) / 96; break;
    case VO_RW_GAMMA: x[0] = 0.5 + 0.25 * (double)(i % 13); break;
    case VO_RW_BETA: x[0] = 0.05 + 0.9 * (double)(i % 17) / 16; break;
    case VO_RW_NEG_GAMMA: x[0] = -0.5 - 0.3 * (double)(i % 9); break;
    case VO_RW_MV3: case VO_RW_BLOCK: case VO_RW_BLOCK1:
        x[0] = 0.1 * (double)(i % 5) - 0.2;
        x[1] = 0.5;
        x[2] = 1 + 0.1 * (double)(i % 7);
        break;
    case VO_RW_FLAT: x[0] = 0; break;
    case VO_RW_FLAT_MV: x[0] = 1; x[1] = 2; x[2] = 3; break;
    case VO_RW_G: x[0] = 0.3 * (double)(i % 7); break;
    default: break;
    }
}

/* One particle of a scenario: `iters` calls of the move, each doing
 * `steps` updates as the scenario's operator form with a proposal object
 * built at the start of the call. out[VO_RW_OUT]:
 * x[0..2], log gamma(x), accepted, g[0..1], words used. */
static void rw_particle(int scenario, uint32_t key, size_t i, size_t iters,
    size_t steps, double *out)
{
    rw_eng g;
    vo_philox_seed(&g.e, key);
    g.used = 0;
    double x[RW_MAXD] = {0, 0, 0, 0};
    double gv[2] = {0, 0};
    vo_rw_init(scenario, i, x);
    const size_t m = (size_t)vo_rw_dim(scenario);
    rw_walk w = {m, 0};
    rw_prop p;
    rw_target lt = lt_flat;
    double lt_x = 0;
    int acc = 0;
    double sd = 1, a1 = -INFINITY, b1 = INFINITY;
    switch (scenario) {
    case VO_RW_NORMAL: sd = 1.5; lt = lt_normal; break;
    case VO_RW_GAMMA: sd = 0.7; a1 = 0; lt = lt_gamma; break;
    case VO_RW_BETA: sd = 0.9; a1 = 0; b1 = 1; lt = lt_beta; break;
    case VO_RW_NEG_GAMMA: sd = 0.6; b1 = 0; lt = lt_neg_gamma; break;
    case VO_RW_MV3: lt = lt_mv3; break;
    case VO_RW_BLOCK: lt = lt_mv3; w.dim = 2; break;
    case VO_RW_BLOCK1: sd = 0.5; a1 = 0; lt = lt_mv3; w.dim = 1; break;
    case VO_RW_FLAT: sd = 1.3; lt = lt_flat; break;
    case VO_RW_FLAT_MV: lt = lt_flat; break;
    case VO_RW_G: sd = 1.1; lt = lt_normal; w.dim_g = 2; break;
    default: return;
    }
    if (scenario == VO_RW_G) {
        gv[0] = x[0];
        gv[1] = x[0] * x[0];
    }
    lt_x = lt(m, x, NULL);
    for (size_t it = 0; it < iters; ++it) {
        /* each move call builds its proposal afresh (a functor's local
         * object), so a cached Box-Muller half does not outlive the call */
        if (scenario == VO_RW_MV3)
            prop_mv(&p, 3, kChol3, kA3, kB3);
        else if (scenario == VO_RW_BLOCK)
            prop_mv(&p, 2, kChol2, kA2, kB2);
        else if (scenario == VO_RW_FLAT_MV)
            prop_mv(&p, 3, kChol3, NULL, NULL);
        else
            prop_normal(&p, sd, a1, b1);
        switch (scenario) {
        case VO_RW_NORMAL: case VO_RW_FLAT_MV: /* one-step, ltx kept */
            for (size_t k = 0; k < steps; ++k)
                acc += rw_step(&w, &g, m, 0, x, &lt_x, NULL, lt, &p);
            break;
        case VO_RW_BETA: case VO_RW_FLAT: /* one-step, ltx null */
            for (size_t k = 0; k < steps; ++k)
                acc += rw_step(&w, &g, m, 0, x, NULL, NULL, lt, &p);
            break;
        case VO_RW_GAMMA: case VO_RW_MV3: /* n-step, ltx kept */
            acc += rw_steps(&w, steps, &g, m, 0, x, &lt_x, NULL, lt, &p);
            break;
        case VO_RW_NEG_GAMMA: /* n-step, ltx null */
            acc += rw_steps(&w, steps, &g, m, 0, x, NULL, NULL, lt, &p);
            break;
        case VO_RW_BLOCK: /* n-step block of x[1..3), ltx kept */
            acc += rw_steps(&w, steps, &g, m, 1, x, &lt_x, NULL, lt, &p);
            break;
        case VO_RW_BLOCK1: /* one-step block of x[2], ltx null */
            for (size_t k = 0; k < steps; ++k)
                acc += rw_step(&w, &g, m, 2, x, NULL, NULL, lt, &p);
            break;
        case VO_RW_G: /* RandomWalkG one-step, ltx and g kept */
            for (size_t k = 0; k < steps; ++k)
                acc += rw_step(&w, &g, m, 0, x, &lt_x, gv, lt, &p);
            break;
        default: break;
        }
    }
    /* scenarios with ltx null report log gamma of the final state */
    if (scenario == VO_RW_BETA || scenario == VO_RW_FLAT
        || scenario == VO_RW_NEG_GAMMA || scenario == VO_RW_BLOCK1)
        lt_x = lt(m, x, NULL);
    out[0] = x[0];
    out[1] = m > 1 ? x[1] : 0;
    out[2] = m > 2 ? x[2] : 0;
    out[3] = lt_x;
    out[4] = (double)acc;
    out[5] = gv[0];
    out[6] = gv[1];
    out[7] = (double)g.used;
}

int vo_rw_run(int scenario, uint32_t seed, size_t n, size_t iters,
    size_t steps, double *out)
{
    if (vo_rw_dim(scenario) == 0)
        return -1;
    /* particle i: key seed_key(seed, i), the i-th seed Seed::get() hands out
     * after Seed::set(seed) (include/vsmc/rng/seed.hpp:118-123) */
    const uint64_t m = 0xFFFFFFFFull;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const size_t i = (size_t)ii;
        const uint32_t key = (uint32_t)(((uint64_t)seed % m + i) % m) + 1u;
        rw_particle(scenario, key, i, iters, steps, out + i * VO_RW_OUT);
    }
    return 0;
}
