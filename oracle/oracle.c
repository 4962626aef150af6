/*
 * oracle.c -- CPU restatement of vSMC's per-step particle path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP path. It is
 * compiled by oracle/Makefile into oracle/liboracle.so and loaded by tests/,
 * the smoke check and the CPU-baseline leg of bench.py; the product library
 * never links or calls it.
 *
 * How each part is pinned (what makes a plausible mistake fail a test):
 *   vo_philox4x32_10, vo_threefry4x64_20, engines, u01 conversions
 *       -- bit-equal to the reference's own C restatement
 *          include/vsmc/rngc/{philox,threefry,u01}.h compiled unchanged from
 *          /root/reference into oracle/_ref (tests/test_oracle_rng.py), and
 *          the committed word fixtures in tests/golden/rng_words.json.
 *   vo_philox_fill / discard -- against repeated vo_philox_next (the
 *          reference's loop == batch self-test, example/rng/include/
 *          rng_test.hpp:80-96).
 *   normals -- closed form Box-Muller from the pinned U01 values (the test
 *          recomputes r*cos / r*sin with numpy), moments and KS statistic.
 *   weights -- closed form softmax, sum W = 1, 1 <= ESS <= N, T = 1
 *          log-evidence closed form.  ddot order: PARITY UNPINNED (the
 *          reference calls an unvendored CBLAS ddot; a stated FP64
 *          tolerance is used).
 *   trans_u01_rep / rep_index / u01_index / index_rep / residual
 *          -- brute force on tiny inputs (every u against an explicit
 *          cumulative table built from a Python left fold), invariants
 *          (sum rep = N, idx[i] = i when rep[i] > 0), M in {0,1,2}, N = 0,
 *          u == cumsum ties.
 *   schemes -- the u01 sequences against their closed forms; the number of
 *          engine words consumed against the lazy-draw rule.
 *   copy, monitor -- closed forms.
 *   PF -- T = 1 log-evidence closed form; linear-Gaussian log-evidence
 *          against vo_kalman_cv; Kalman against a brute-force dense
 *          Gaussian marginal likelihood on T <= 3 (numpy).
 *   random walk (rw.c) -- flat-target chains against the stream's normals
 *          built from raw words, transform closed forms and finite-
 *          difference Jacobians, stationary moments of long chains, word
 *          counts; the Cholesky summation order is PARITY UNPINNED
 *          (unvendored CBLAS tpmv in the reference).
 *
 * Floating-point: built with -ffp-contract=off (the reference's x86-64
 * build has no FMA, and u01_sequence.hpp:200,:246 must not be contracted).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ RNG */

/* One Philox4x32 round: include/vsmc/rng/philox.hpp:208-228. The products
 * are 32x32->64; state word 2 uses the second multiplier. */
static void philox_round(uint32_t s[4], const uint32_t par[2])
{
    uint64_t p0 = (uint64_t)s[0] * UINT64_C(0xD2511F53);
    uint64_t p2 = (uint64_t)s[2] * UINT64_C(0xCD9E8D57);
    uint32_t hi0 = (uint32_t)(p2 >> 32) ^ par[0];
    uint32_t lo1 = (uint32_t)p2;
    uint32_t hi2 = (uint32_t)(p0 >> 32) ^ par[1];
    uint32_t lo3 = (uint32_t)p0;
    s[0] = hi0 ^ s[1];
    s[1] = lo1;
    s[2] = hi2 ^ s[3];
    s[3] = lo3;
}

/* include/vsmc/rng/philox.hpp:263-295: round 1 with the raw key, then the
 * key is bumped by the Weyl constants (:76-80) before each of rounds 2-10. */
void vo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
    uint32_t out[4])
{
    uint32_t s[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t par[2] = {key[0], key[1]};
    philox_round(s, par);
    for (int r = 2; r <= 10; ++r) {
        par[0] += UINT32_C(0x9E3779B9);
        par[1] += UINT32_C(0xBB67AE85);
        philox_round(s, par);
    }
    memcpy(out, s, sizeof(s));
}

static uint64_t rotl64(uint64_t x, int r)
{
    return (x << r) | (x >> (64 - r));
}

/* Rotation constants R[(N-1) % 8][0|1], include/vsmc/rng/threefry.hpp:
 * 120-135 (4x64 rows). */
static const int tf_rot[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
    {25, 33}, {46, 12}, {58, 22}, {32, 32}};

/* include/vsmc/rng/threefry.hpp:148-298: key schedule par[0..3] = key,
 * par[4] = 0x1BD11BDAA9FC1A22 ^ key0 ^ .. ^ key3; key injection before
 * round 1 and after every 4th round (inc = N/4); round N mixes (0, i0) and
 * (2, i2) with i0 = 1, i2 = 3 for odd N and i0 = 3, i2 = 1 for even N. */
void vo_threefry4x64_20(const uint64_t ctr[4], const uint64_t key[4],
    uint64_t out[4])
{
    uint64_t par[5];
    uint64_t s[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    par[4] = UINT64_C(0x1BD11BDAA9FC1A22);
    for (int k = 0; k < 4; ++k) {
        par[k] = key[k];
        par[4] ^= key[k];
    }
    for (int k = 0; k < 4; ++k)
        s[k] += par[k % 5];
    for (int n = 1; n <= 20; ++n) {
        const int i0 = (n % 2) ? 1 : 3;
        const int i2 = (n % 2) ? 3 : 1;
        const int r = (n - 1) % 8;
        s[0] += s[i0];
        s[i0] = rotl64(s[i0], tf_rot[r][0]);
        s[i0] ^= s[0];
        s[2] += s[i2];
        s[i2] = rotl64(s[i2], tf_rot[r][1]);
        s[i2] ^= s[2];
        if (n % 4 == 0) {
            const int inc = n / 4;
            for (int k = 0; k < 4; ++k)
                s[k] += par[(inc + k) % 5];
            s[3] += (uint64_t)inc;
        }
    }
    memcpy(out, s, sizeof(s));
}

/* increment(ctr): include/vsmc/rng/counter.hpp:42-72 (carry word by word) */
static void inc_ctr32(uint32_t c[4])
{
    for (int k = 0; k < 4; ++k)
        if (++c[k] != 0)
            return;
}

static void inc_ctr64(uint64_t c[4])
{
    for (int k = 0; k < 4; ++k)
        if (++c[k] != 0)
            return;
}

/* increment(ctr, nskip): include/vsmc/rng/counter.hpp:74-93. The carry is
 * taken whenever front >= max - nskip, including front + nskip == max. */
static void inc_ctr32_n(uint32_t c[4], uint32_t nskip)
{
    if (c[0] < UINT32_MAX - nskip) {
        c[0] += nskip;
    } else {
        c[0] += nskip;
        for (int k = 1; k < 4; ++k)
            if (++c[k] != 0)
                break;
    }
}

static void inc_ctr64_n(uint64_t c[4], uint64_t nskip)
{
    if (c[0] < UINT64_MAX - nskip) {
        c[0] += nskip;
    } else {
        c[0] += nskip;
        for (int k = 1; k < 4; ++k)
            if (++c[k] != 0)
                break;
    }
}

/* CounterEngine::seed(s): key = {s, 0, ...}, ctr = 0, index = 4
 * (include/vsmc/rng/counter.hpp:309-314, :490-495) */
void vo_philox_seed(vo_philox *e, uint32_t s)
{
    memset(e, 0, sizeof(*e));
    e->key[0] = s;
    e->index = 4;
}

/* CounterEngine::operator()(): increment then encrypt when the buffer is
 * used up (include/vsmc/rng/counter.hpp:347-355, philox.hpp:263-269) */
uint32_t vo_philox_next(vo_philox *e)
{
    if (e->index == 4) {
        inc_ctr32(e->ctr);
        vo_philox4x32_10(e->ctr, e->key, e->buf);
        e->index = 0;
    }
    return e->buf[e->index++];
}

void vo_philox_set_ctr(vo_philox *e, const uint32_t ctr[4])
{
    memcpy(e->ctr, ctr, sizeof(e->ctr));
    e->index = 4;
}

/* generator_(ctr, key, cnt, blocks): block counters as increment(ctr, n,
 * block) lays them out (include/vsmc/rng/counter.hpp:181-215): ctr + 1,
 * then + q on the low word, with the carry rule of increment(ctr, nskip). */
static void philox_gen_blocks(vo_philox *e, size_t cnt, uint32_t *out)
{
    if (cnt == 0)
        return;
    uint32_t c[4];
    memcpy(c, e->ctr, sizeof(c));
    inc_ctr32(c);
    const uint64_t m = UINT32_MAX;
    const uint64_t l = c[0];
    const uint64_t k = cnt;
    uint32_t blk[4];
    for (size_t q = 0; q != cnt; ++q) {
        memcpy(blk, c, sizeof(blk));
        if (k < m && l < m - k)
            blk[0] += (uint32_t)q;
        else
            inc_ctr32_n(blk, (uint32_t)q);
        vo_philox4x32_10(blk, e->key, out + 4 * q);
    }
    memcpy(e->ctr, blk, sizeof(blk));
}

/* CounterEngine::operator()(n, r), include/vsmc/rng/counter.hpp:357-390:
 * drain the buffer, then chunks of 256 blocks, then the leftover blocks,
 * then single words. */
void vo_philox_fill(vo_philox *e, size_t n, uint32_t *r)
{
    const size_t remain = 4 - e->index;
    if (n < remain) {
        memcpy(r, e->buf + e->index, sizeof(uint32_t) * n);
        e->index += (uint32_t)n;
        return;
    }
    memcpy(r, e->buf + e->index, sizeof(uint32_t) * remain);
    r += remain;
    n -= remain;
    e->index = 4;
    const size_t kb = 1024 / 4;
    const size_t mchunks = (n / 4) / kb;
    const size_t lblocks = (n / 4) % kb;
    for (size_t i = 0; i != mchunks; ++i) {
        philox_gen_blocks(e, kb, r);
        r += 4 * kb;
        n -= 4 * kb;
    }
    philox_gen_blocks(e, lblocks, r);
    r += 4 * lblocks;
    n -= 4 * lblocks;
    for (size_t i = 0; i != n; ++i)
        r[i] = vo_philox_next(e);
}

/* CounterEngine::discard, include/vsmc/rng/counter.hpp:392-412 */
void vo_philox_discard(vo_philox *e, uint32_t nskip)
{
    size_t n = nskip;
    if (e->index + n <= 4) {
        e->index += (uint32_t)n;
        return;
    }
    n -= 4 - e->index;
    if (n <= 4) {
        e->index = 4;
        vo_philox_next(e);
        e->index = (uint32_t)n;
        return;
    }
    inc_ctr32_n(e->ctr, (uint32_t)(n / 4));
    e->index = 4;
    vo_philox_next(e);
    e->index = (uint32_t)(n % 4);
}

void vo_threefry_seed(vo_threefry *e, uint64_t s)
{
    memset(e, 0, sizeof(*e));
    e->key[0] = s;
    e->index = 4;
}

uint64_t vo_threefry_next(vo_threefry *e)
{
    if (e->index == 4) {
        inc_ctr64(e->ctr);
        vo_threefry4x64_20(e->ctr, e->key, e->buf);
        e->index = 0;
    }
    return e->buf[e->index++];
}

void vo_threefry_set_ctr(vo_threefry *e, const uint64_t ctr[4])
{
    memcpy(e->ctr, ctr, sizeof(e->ctr));
    e->index = 4;
}

void vo_threefry_fill(vo_threefry *e, size_t n, uint64_t *r)
{
    const size_t remain = 4 - e->index;
    if (n < remain) {
        memcpy(r, e->buf + e->index, sizeof(uint64_t) * n);
        e->index += (uint32_t)n;
        return;
    }
    memcpy(r, e->buf + e->index, sizeof(uint64_t) * remain);
    r += remain;
    n -= remain;
    e->index = 4;
    const size_t kb = 1024 / 4;
    const size_t nblocks = n / 4;
    size_t done = 0;
    while (done < nblocks) {
        const size_t cnt = (nblocks - done) < kb ? (nblocks - done) : kb;
        uint64_t c[4];
        memcpy(c, e->ctr, sizeof(c));
        inc_ctr64(c);
        const uint64_t m = UINT64_MAX;
        const uint64_t l = c[0];
        const uint64_t k = cnt;
        uint64_t blk[4];
        for (size_t q = 0; q != cnt; ++q) {
            memcpy(blk, c, sizeof(blk));
            if (k < m && l < m - k)
                blk[0] += (uint64_t)q;
            else
                inc_ctr64_n(blk, (uint64_t)q);
            vo_threefry4x64_20(blk, e->key, r + 4 * q);
            if (q + 1 == cnt)
                memcpy(e->ctr, blk, sizeof(blk));
        }
        r += 4 * cnt;
        n -= 4 * cnt;
        done += cnt;
    }
    for (size_t i = 0; i != n; ++i)
        r[i] = vo_threefry_next(e);
}

void vo_threefry_discard(vo_threefry *e, uint64_t nskip)
{
    uint64_t n = nskip;
    if (e->index + n <= 4) {
        e->index += (uint32_t)n;
        return;
    }
    n -= 4 - e->index;
    if (n <= 4) {
        e->index = 4;
        vo_threefry_next(e);
        e->index = (uint32_t)n;
        return;
    }
    inc_ctr64_n(e->ctr, n / 4);
    e->index = 4;
    vo_threefry_next(e);
    e->index = (uint32_t)(n % 4);
}

/* Word k of the stream is word k % 4 of the block with counter k / 4 + 1
 * (pre-increment, counter.hpp:347-355). Generated block by block so that
 * offsets beyond 2^32 words stay exact. */
void vo_philox_words(uint32_t key0, uint64_t offset, size_t n, uint32_t *out)
{
    uint32_t key[2] = {key0, 0};
    uint32_t blk[4];
    uint64_t cur = (uint64_t)-1;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t w = offset + i;
        const uint64_t b = w / 4 + 1;
        if (b != cur) {
            uint32_t ctr[4] = {(uint32_t)b, (uint32_t)(b >> 32), 0, 0};
            vo_philox4x32_10(ctr, key, blk);
            cur = b;
        }
        out[i] = blk[w % 4];
    }
}

void vo_threefry_words(uint64_t key0, uint64_t offset, size_t n,
    uint64_t *out)
{
    uint64_t key[4] = {key0, 0, 0, 0};
    uint64_t blk[4];
    uint64_t cur = (uint64_t)-1;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t w = offset + i;
        const uint64_t b = w / 4 + 1;
        if (b != cur) {
            uint64_t ctr[4] = {b, 0, 0, 0};
            vo_threefry4x64_20(ctr, key, blk);
            cur = b;
        }
        out[i] = blk[w % 4];
    }
}

/* u01_co (U01LRImpl<UInt, Real, Closed, Open>), include/vsmc/rng/u01.hpp:
 * 126-147: P = min(W, digits(Real)), value = (u >> (W - P)) * 2^-P */
double vo_u01_co_u32d(uint32_t u)
{
    return (double)u * (1.0 / 4294967296.0);
}

double vo_u01_co_u64d(uint64_t u)
{
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

float vo_u01_co_u32f(uint32_t u)
{
    return (float)(u >> 8) * (1.0f / 16777216.0f);
}

void vo_u01_fill(int kind, uint64_t key0, uint64_t offset, size_t n,
    double *out)
{
    if (kind == VO_PHILOX4X32) {
        uint32_t *w = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        vo_philox_words((uint32_t)key0, offset, n, w);
        for (size_t i = 0; i < n; ++i)
            out[i] = vo_u01_co_u32d(w[i]);
        free(w);
    } else {
        uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
        vo_threefry_words(key0, offset, n, w);
        for (size_t i = 0; i < n; ++i)
            out[i] = vo_u01_co_u64d(w[i]);
        free(w);
    }
}

/* 2 pi as const_pi_2<double>() (include/vsmc/math/constants.hpp:66) */
static const double two_pi = 6.283185307179586476925286766559005768394;

/* normal_distribution_impl<1024>(rng, n, r, mean, sd), include/vsmc/rng/
 * normal_distribution.hpp:222-241: the first n/2 uniforms are radii, the
 * next n/2 angles; r[i] = sd*sqrt(-2 log u_i) * sin(2 pi u_{n/2+i}) + mean,
 * r[n/2+i] = the same with cos. The odd tail (:259-266) draws two more. */
void vo_normal_fill(int kind, uint64_t key0, uint64_t offset, size_t n,
    double mean, double sd, double *out)
{
    uint64_t pos = offset;
    size_t done = 0;
    double u[1024];
    while (done < n) {
        const size_t k = (n - done) < 1024 ? (n - done) : 1024;
        const size_t nu = k / 2;
        vo_u01_fill(kind, key0, pos, k, u);
        pos += k;
        double *r = out + done;
        for (size_t i = 0; i < nu; ++i) {
            double s = log(u[i]);
            s = -2 * s;
            s = sqrt(s);
            double a = two_pi * u[nu + i];
            double sn = sin(a);
            double cs = cos(a);
            s = sd * s;
            r[i] = s * sn + mean;
            r[nu + i] = s * cs + mean;
        }
        if (k % 2 != 0)
            r[k - 1] = u[k - 1];
        done += k;
        if (done == n && (n % 2) != 0) {
            double uv[2];
            vo_u01_fill(kind, key0, pos, 2, uv);
            pos += 2;
            r[k - 1] = mean + sd * sqrt(-2 * log(uv[0])) * cos(two_pi * uv[1]);
        }
    }
}

/* Scalar NormalDistribution over the stream: pairs of words give
 * (r cos, r sin), the sine half cached (normal_distribution.hpp:199-216). */
void vo_normal_fill_scalar(int kind, uint64_t key0, uint64_t offset,
    size_t n, double mean, double sd, double *out)
{
    const size_t npair = (n + 1) / 2;
    double *u = (double *)malloc(sizeof(double) * 2 * (npair ? npair : 1));
    vo_u01_fill(kind, key0, offset, 2 * npair, u);
    for (size_t p = 0; p < npair; ++p) {
        const double r = sqrt(-2 * log(u[2 * p]));
        const double a = two_pi * u[2 * p + 1];
        const double z = r * cos(a);
        const double v = r * sin(a);
        out[2 * p] = mean + sd * z;
        if (2 * p + 1 < n)
            out[2 * p + 1] = mean + sd * v;
    }
    free(u);
}

void vo_normal_init(vo_normal *d, double mean, double sd)
{
    d->mean = mean;
    d->sd = sd;
    d->v = 0;
    d->saved = 0;
}

double vo_normal_philox(vo_normal *d, vo_philox *e)
{
    double z;
    if (d->saved) {
        z = d->v;
        d->saved = 0;
    } else {
        const double u1 = sqrt(-2 * log(vo_u01_co_u32d(vo_philox_next(e))));
        const double u2 = two_pi * vo_u01_co_u32d(vo_philox_next(e));
        z = u1 * cos(u2);
        d->v = u1 * sin(u2);
        d->saved = 1;
    }
    return d->mean + d->sd * z;
}

/* SeedGenerator: default divisor 1, remainder 0, seed_max = UINT_MAX;
 * get() = skip() then seed * div + rem (include/vsmc/rng/seed.hpp:
 * 111-178, :221-225) */
void vo_seed_init(vo_seedgen *g)
{
    g->seed = 0;
    g->divisor = 1;
    g->remainder = 0;
    g->seed_max = UINT32_MAX;
    g->seed_max -= g->seed_max % g->divisor;
    g->seed_max /= g->divisor;
}

void vo_seed_set(vo_seedgen *g, uint32_t s)
{
    g->seed = s % g->seed_max;
}

uint32_t vo_seed_get(vo_seedgen *g)
{
    if (g->seed < g->seed_max)
        ++g->seed;
    else
        g->seed = 1;
    return g->seed * g->divisor + g->remainder;
}

/* --------------------------------------------------------------- Weight */

/* weight_ess: 1 / ddot(W, W) (include/vsmc/core/weight.hpp:43-47). The
 * ddot summation order of the CBLAS the reference links is not vendored;
 * a sequential left fold is used here (parity unpinned beyond FP64
 * tolerance). */
double vo_weight_ess(size_t n, const double *W)
{
    double acc = 0;
    for (size_t i = 0; i < n; ++i)
        acc += W[i] * W[i];
    return 1 / acc;
}

/* Weight::normalize(use_log), include/vsmc/core/weight.hpp:208-222: exp in
 * chunks of 1024 (when use_log), one sequential accumulate across chunks,
 * then multiply by 1 / accw (mul(N, a, b, y): y = a * b, vmath.hpp:228). */
static double normalize(size_t n, double *W, int use_log, double *accw_out)
{
    double accw = 0;
    for (size_t i = 0; i < n; ++i) {
        if (use_log)
            W[i] = exp(W[i]);
        accw += W[i];
    }
    const double inv = 1 / accw;
    for (size_t i = 0; i < n; ++i)
        W[i] = inv * W[i];
    if (accw_out)
        *accw_out = accw;
    return vo_weight_ess(n, W);
}

/* Weight::set_equal (weight.hpp:99-103) */
double vo_weight_set_equal(size_t n, double *W)
{
    for (size_t i = 0; i < n; ++i)
        W[i] = 1.0 / (double)n;
    return normalize(n, W, 0, NULL);
}

/* post_set_log: subtract max (weight_normalize_log, weight.hpp:58-63), then
 * normalize(true). logz_inc = wmax + log(accw). */
static double post_set_log(size_t n, double *W, double *logz_inc)
{
    double wmax = W[0];
    for (size_t i = 1; i < n; ++i)
        if (W[i] > wmax)
            wmax = W[i];
    for (size_t i = 0; i < n; ++i)
        W[i] -= wmax;
    double accw;
    const double ess = normalize(n, W, 1, &accw);
    if (logz_inc)
        *logz_inc = wmax + log(accw);
    return ess;
}

/* Weight::set_log (weight.hpp:144-149). The log-evidence increment assumes
 * the previous weights were equal (1/N), as after do_reset / set_equal. */
double vo_weight_set_log(size_t n, double *W, const double *logw,
    double *logz_inc)
{
    memcpy(W, logw, sizeof(double) * n);
    const double ess = post_set_log(n, W, logz_inc);
    if (logz_inc)
        *logz_inc -= log((double)n);
    return ess;
}

/* Weight::add_log (weight.hpp:168-173): W <- log(W) + inc, post_set_log.
 * Because W is normalised, max + log sum exp(log W + inc - max) is
 * log sum W_i exp(inc_i), the evidence increment. */
double vo_weight_add_log(size_t n, double *W, const double *inc,
    double *logz_inc)
{
    for (size_t i = 0; i < n; ++i)
        W[i] = log(W[i]);
    for (size_t i = 0; i < n; ++i)
        W[i] = W[i] + inc[i];
    return post_set_log(n, W, logz_inc);
}

/* ------------------------------------------------------------- Resample */

/* u01_sorted: include/vsmc/rng/u01_sequence.hpp:53-62 */
void vo_u01_sorted(size_t N, const double *u01, double *out)
{
    for (size_t i = 0; i < N; ++i)
        out[i] = log(u01[i]);
    double lmax = 0;
    for (size_t i = 0; i < N; ++i) {
        lmax += out[i] / (double)(N - i);
        out[i] = 1 - exp(lmax);
    }
}

/* u01_stratified: u01_sequence.hpp:64-71 */
void vo_u01_stratified(size_t N, const double *u01, double *out)
{
    const double delta = 1 / (double)N;
    for (size_t i = 0; i < N; ++i)
        out[i] = u01[i] * delta + (double)i * delta;
}

/* u01_systematic (the cumulative free-function form, u01_sequence.hpp:
 * 73-89; the sampler's U01SequenceSystematic uses u0 + j * delta instead) */
void vo_u01_systematic(size_t N, double u01, double *out)
{
    if (N == 0)
        return;
    const double delta = 1.0 / (double)N;
    out[0] = u01 * delta;
    for (size_t i = 1; i < N; ++i)
        out[i] = out[i - 1] + delta;
}

/* Lazy u01 sequences of the schemes (u01_sequence.hpp:130-259). */
typedef struct {
    int kind;      /* 0 sorted, 1 stratified, 2 systematic, 3 array */
    size_t N;
    size_t n;      /* last index evaluated (N = none yet) */
    double u;
    double lmax;
    double delta;
    double u0;
    vo_philox *rng;
    const double *arr;
    uint64_t words;
} useq;

static void useq_init(useq *q, int kind, size_t N, vo_philox *rng,
    const double *arr)
{
    memset(q, 0, sizeof(*q));
    q->kind = kind;
    q->N = N;
    q->n = N;
    q->rng = rng;
    q->arr = arr;
    q->delta = 1 / (double)N;
    if (kind == 2) {
        /* U01SequenceSystematic draws u0 in its constructor (:231-236) */
        q->u0 = vo_u01_co_u32d(vo_philox_next(rng)) * q->delta;
        q->words = 1;
    }
}

static double useq_at(useq *q, size_t n)
{
    if (q->kind == 3)
        return q->arr[n];
    if (n == q->n)
        return q->u;
    if (q->kind == 0) {
        /* U01SequenceSorted::operator[] (:141-154) */
        const double u01 = vo_u01_co_u32d(vo_philox_next(q->rng));
        q->words++;
        q->lmax += log(u01) / (double)(q->N - n);
        q->n = n;
        q->u = 1 - exp(q->lmax);
    } else if (q->kind == 1) {
        /* U01SequenceStratified::operator[] (:191-203) */
        const double u01 = vo_u01_co_u32d(vo_philox_next(q->rng));
        q->words++;
        q->n = n;
        q->u = u01 * q->delta + (double)n * q->delta;
    } else {
        /* U01SequenceSystematic::operator[] (:240-249) */
        q->n = n;
        q->u = q->u0 + (double)n * q->delta;
    }
    return q->u;
}

/* resample_trans_u01_rep, include/vsmc/resample/transform.hpp:42-71 */
static void trans_u01_rep_seq(size_t M, size_t N, const double *W, useq *q,
    int64_t *rep)
{
    if (M == 0)
        return;
    if (M == 1) {
        rep[0] = (int64_t)N;
        return;
    }
    memset(rep, 0, sizeof(int64_t) * M);
    if (N == 0)
        return;
    double accw = 0;
    size_t j = 0;
    for (size_t i = 0; i != M - 1; ++i) {
        accw += W[i];
        while (j != N && useq_at(q, j) <= accw) {
            ++rep[i];
            ++j;
        }
    }
    rep[M - 1] = (int64_t)(N - j);
}

void vo_trans_u01_rep(size_t M, size_t N, const double *W, const double *u,
    int64_t *rep)
{
    useq q;
    useq_init(&q, 3, N, NULL, u);
    trans_u01_rep_seq(M, N, W, &q, rep);
}

/* resample_trans_u01_index, transform.hpp:75-95 */
void vo_trans_u01_index(size_t M, size_t N, const double *W, const double *u,
    int64_t *idx)
{
    if (M == 0 || N == 0)
        return;
    memset(idx, 0, sizeof(int64_t) * N);
    if (M == 1)
        return;
    double accw = 0;
    size_t j = 0;
    for (size_t i = 0; i != M - 1; ++i) {
        accw += W[i];
        while (j != N && u[j] <= accw)
            idx[j++] = (int64_t)i;
    }
    while (j != N)
        idx[j++] = (int64_t)(M - 1);
}

/* resample_trans_rep_index, transform.hpp:99-134. For M == N a particle
 * with children keeps its own slot; the k-th childless slot (in dst order)
 * takes the source whose extra copies (rep - 1, in src order) cover k. */
void vo_trans_rep_index(size_t M, size_t N, const int64_t *rep,
    int64_t *idx)
{
    if (M == 0 || N == 0)
        return;
    if (M != N) {
        size_t dst = 0;
        for (size_t src = 0; src != M; ++src)
            for (int64_t r = 0; r != rep[src]; ++r)
                idx[dst++] = (int64_t)src;
        return;
    }
    int64_t time = 0;
    size_t src = 0;
    for (size_t dst = 0; dst != N; ++dst) {
        if (rep[dst] != 0) {
            idx[dst] = (int64_t)dst;
        } else {
            if (rep[src] < time + 2) {
                time = 0;
                do
                    ++src;
                while (rep[src] < 2);
            }
            idx[dst] = (int64_t)src;
            ++time;
        }
    }
}

/* resample_trans_index_rep, transform.hpp:138-149 */
void vo_trans_index_rep(size_t M, size_t N, const int64_t *idx,
    int64_t *rep)
{
    if (M == 0 || N == 0)
        return;
    memset(rep, 0, sizeof(int64_t) * M);
    for (size_t i = 0; i != N; ++i)
        ++rep[idx[i]];
}

/* ResampleIndex::index, include/vsmc/resample/index.hpp:121-134: start
 * from the last row at id, then apply the rows before it down to iter. */
int64_t vo_rindex_index(size_t N, size_t T, const int64_t *hist, size_t id,
    size_t iter)
{
    size_t cur = T - 1;
    int64_t idx = hist[cur * N + id];
    while (cur != iter) {
        --cur;
        idx = hist[cur * N + (size_t)idx];
    }
    return idx;
}

/* ResampleIndex::index_matrix_dispatch, index.hpp:156-204: the last
 * iteration's column is the last row; each earlier one maps the next one
 * through that iteration's row. */
void vo_rindex_matrix(size_t N, size_t T, const int64_t *hist, int layout,
    int64_t *out)
{
    if (N * T == 0)
        return;
    const int row_major = layout == 101;
    for (size_t i = 0; i != N; ++i)
        out[row_major ? i * T + (T - 1) : (T - 1) * N + i] = hist[(T - 1) * N + i];
    for (size_t t = T - 1; t != 0; --t) {
        const int64_t *row = hist + (t - 1) * N;
        for (size_t i = 0; i != N; ++i) {
            const int64_t last = out[row_major ? i * T + t : t * N + i];
            out[row_major ? i * T + (t - 1) : (t - 1) * N + i] = row[(size_t)last];
        }
    }
}

/* resample_trans_residual, transform.hpp:154-168: resid_i, integ_i =
 * modf(N * W_i); resid *= 1 / (sequential sum of resid); returns
 * N - sum(integ). */
size_t vo_trans_residual(size_t M, size_t N, const double *W, double *resid,
    int64_t *integ)
{
    double integral = 0;
    const double coeff = (double)N;
    for (size_t i = 0; i != M; ++i) {
        resid[i] = modf(coeff * W[i], &integral);
        integ[i] = (int64_t)integral;
    }
    double acc = 0;
    for (size_t i = 0; i != M; ++i)
        acc += resid[i];
    const double inv = 1 / acc;
    for (size_t i = 0; i != M; ++i)
        resid[i] = inv * resid[i];
    int64_t R = 0;
    for (size_t i = 0; i != M; ++i)
        R += integ[i];
    return N - (size_t)R;
}

/* Resample{Multinomial,Stratified,Systematic} (multinomial.hpp:43-53,
 * stratified.hpp:43-53, systematic.hpp:43-53) and the residual variants
 * (residual.hpp:43-58, residual_stratified.hpp, residual_systematic.hpp). */
uint64_t vo_resample(int scheme, size_t M, size_t N, vo_philox *rng,
    const double *W, int64_t *rep)
{
    useq q;
    if (scheme == VO_MULTINOMIAL || scheme == VO_STRATIFIED ||
        scheme == VO_SYSTEMATIC) {
        const int kind = scheme == VO_MULTINOMIAL ? 0 :
            (scheme == VO_STRATIFIED ? 1 : 2);
        useq_init(&q, kind, N, rng, NULL);
        trans_u01_rep_seq(M, N, W, &q, rep);
        return q.words;
    }
    double *resid = (double *)malloc(sizeof(double) * (M ? M : 1));
    int64_t *integ = (int64_t *)malloc(sizeof(int64_t) * (M ? M : 1));
    const size_t R = vo_trans_residual(M, N, W, resid, integ);
    const int kind = scheme == VO_RESIDUAL ? 0 :
        (scheme == VO_RESIDUAL_STRATIFIED ? 1 : 2);
    useq_init(&q, kind, R, rng, NULL);
    trans_u01_rep_seq(M, R, resid, &q, rep);
    for (size_t i = 0; i != M; ++i)
        rep[i] += integ[i];
    free(resid);
    free(integ);
    return q.words;
}

void vo_scheme_u01(int scheme, size_t N, vo_philox *rng, size_t nu,
    double *out)
{
    useq q;
    const int kind = (scheme == VO_MULTINOMIAL || scheme == VO_RESIDUAL) ? 0 :
        ((scheme == VO_STRATIFIED || scheme == VO_RESIDUAL_STRATIFIED) ? 1 :
                                                                         2);
    useq_init(&q, kind, N, rng, NULL);
    for (size_t j = 0; j < nu; ++j)
        out[j] = useq_at(&q, j);
}

/* StateMatrix<RowMajor>::copy: dst order, skip src == dst
 * (state_matrix.hpp:246-262) */
void vo_copy_rowmajor(size_t N, size_t dim, double *x, const int64_t *idx)
{
    for (size_t dst = 0; dst != N; ++dst) {
        const size_t src = (size_t)idx[dst];
        if (src == dst)
            continue;
        memcpy(x + dst * dim, x + src * dim, sizeof(double) * dim);
    }
}

/* StateMatrix<ColMajor>::copy: per dimension, dst order
 * (state_matrix.hpp:356-364) */
void vo_copy_colmajor(size_t N, size_t dim, double *x, const int64_t *idx)
{
    for (size_t d = 0; d != dim; ++d)
        for (size_t dst = 0; dst != N; ++dst)
            x[d * N + dst] = x[d * N + (size_t)idx[dst]];
}

/* dgemv(ColMajor, NoTrans, dim, N, 1, buf, dim, W, 1, 0, r, 1): r_d =
 * sum_i buf[i*dim+d] * W_i. Summation order of the linked CBLAS is not
 * vendored (parity unpinned beyond FP64 tolerance); left fold here. */
void vo_monitor_moments(size_t N, size_t dim, const double *buf,
    const double *W, double *r)
{
    for (size_t d = 0; d < dim; ++d) {
        double acc = 0;
        for (size_t i = 0; i < N; ++i)
            acc += buf[i * dim + d] * W[i];
        r[d] = acc;
    }
}

/* ------------------------------------------------------------------ PF */

/* log-likelihood of example/pf/include/pf.hpp:78-87 (Student-t, nu = df,
 * scale s: the reference multiplies by 1/s = 10), optionally with the
 * normalising constants the reference drops; or a Gaussian. */
static double pf_loglik(const vo_pf_config *c, double x, double y, double ox,
    double oy)
{
    if (c->obs_model == 2) {
        const double rx = (x - ox) / c->obs_scale;
        const double ry = (y - oy) / c->obs_scale;
        return -0.5 * (rx * rx + ry * ry) -
            2 * (log(c->obs_scale) + 0.5 * log(two_pi));
    }
    const double scale = 1 / c->obs_scale;
    const double nu = c->obs_df;
    double lx = scale * (x - ox);
    double ly = scale * (y - oy);
    lx = log(1 + lx * lx / nu);
    ly = log(1 + ly * ly / nu);
    double ll = -0.5 * (nu + 1) * (lx + ly);
    if (c->obs_model == 1) {
        const double k = lgamma(0.5 * (nu + 1)) - lgamma(0.5 * nu) -
            0.5 * log(nu * 3.141592653589793238462643383279502884197) -
            log(c->obs_scale);
        ll += 2 * k;
    }
    return ll;
}

double vo_cv_loglik(double x, double y, double ox, double oy, int obs_model,
    double obs_scale, double obs_df)
{
    vo_pf_config c;
    memset(&c, 0, sizeof(c));
    c.obs_model = obs_model;
    c.obs_scale = obs_scale;
    c.obs_df = obs_df;
    return pf_loglik(&c, x, y, ox, oy);
}

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Seconds of the last vo_pf_run per phase: 0 the init / move functor with
 * the log-likelihood (the parallel eval_sp loop), 1 the weight update
 * (set_log / add_log: normalisation and ESS), 2 the resampling counts and
 * rep -> index, 3 the state copy, 4 the monitor. */
static double g_pf_phase[5];

void vo_pf_phase_seconds(double out[5])
{
    for (int k = 0; k < 5; ++k)
        out[k] = g_pf_phase[k];
}

/* Sampler<PFState>(N, scheme, threshold) with init PFInit and move PFMove
 * (user_guide/cpp/pf_seq.cpp:70-122, vSMC NormalDistribution), monitor
 * "pos" (dim 2, MonitorMCMC), driven as initialize(obs) + iterate(T - 1)
 * (sampler.hpp:443-476, :735-785). Engines: Seed::set(seed); particle i
 * has key seed + 1 + i, the resampling engine seed + N + 1
 * (rng_set.hpp:88-115, particle.hpp:60-67). */
double vo_pf_run(const vo_pf_config *cfg, const double *obs_x,
    const double *obs_y, double *ess_out, int64_t *resampled_out,
    double *pos_out, double *logz_out, double *state_out, double *W_out)
{
    const size_t N = (size_t)cfg->N;
    const size_t T = (size_t)cfg->T;
    vo_seedgen sg;
    vo_seed_init(&sg);
    vo_seed_set(&sg, cfg->seed);
    vo_philox *eng = (vo_philox *)malloc(sizeof(vo_philox) * N);
    for (size_t i = 0; i < N; ++i)
        vo_philox_seed(&eng[i], vo_seed_get(&sg));
    vo_philox rng;
    vo_philox_seed(&rng, vo_seed_get(&sg));

    double *x = (double *)malloc(sizeof(double) * 4 * N);
    double *W = (double *)malloc(sizeof(double) * N);
    double *w = (double *)malloc(sizeof(double) * N);
    double *buf = (double *)malloc(sizeof(double) * 2 * N);
    int64_t *rep = (int64_t *)malloc(sizeof(int64_t) * N);
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * N);
#ifdef _OPENMP
    const int nthreads = cfg->threads > 0 ? cfg->threads : omp_get_max_threads();
#endif
    for (int k = 0; k < 5; ++k)
        g_pf_phase[k] = 0;
    double tp = now_seconds();
#define VO_PHASE(k)                            \
    do {                                       \
        const double tn_ = now_seconds();      \
        g_pf_phase[k] += tn_ - tp;             \
        tp = tn_;                              \
    } while (0)
    const double t0 = now_seconds();
    /* do_reset: weight().set_equal() */
    double ess = vo_weight_set_equal(N, W);
    double logz = 0;
    for (size_t t = 0; t < T; ++t) {
        if (t == 0) {
            /* PFInit::eval_sp: NormalDistribution(0, sd_pos0) twice (one
             * pair), NormalDistribution(0, sd_vel0) twice (one pair) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
            for (int64_t ii = 0; ii < (int64_t)N; ++ii) {
                const size_t i = (size_t)ii;
                vo_normal np, nv;
                vo_normal_init(&np, 0, cfg->sd_pos0);
                vo_normal_init(&nv, 0, cfg->sd_vel0);
                double *xi = x + 4 * i;
                xi[0] = vo_normal_philox(&np, &eng[i]);
                xi[1] = vo_normal_philox(&np, &eng[i]);
                xi[2] = vo_normal_philox(&nv, &eng[i]);
                xi[3] = vo_normal_philox(&nv, &eng[i]);
                w[i] = pf_loglik(cfg, xi[0], xi[1], obs_x[0], obs_y[0]);
            }
            VO_PHASE(0);
            double dz;
            ess = vo_weight_set_log(N, W, w, &dz);
            logz = dz;
            VO_PHASE(1);
        } else {
            /* PFMove::eval_sp: x += n + dt * vx; y += n + dt * vy;
             * vx += n; vy += n (velocity read before its update) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
            for (int64_t ii = 0; ii < (int64_t)N; ++ii) {
                const size_t i = (size_t)ii;
                vo_normal np, nv;
                vo_normal_init(&np, 0, cfg->sd_pos);
                vo_normal_init(&nv, 0, cfg->sd_vel);
                double *xi = x + 4 * i;
                xi[0] += vo_normal_philox(&np, &eng[i]) + cfg->dt * xi[2];
                xi[1] += vo_normal_philox(&np, &eng[i]) + cfg->dt * xi[3];
                xi[2] += vo_normal_philox(&nv, &eng[i]);
                xi[3] += vo_normal_philox(&nv, &eng[i]);
                w[i] = pf_loglik(cfg, xi[0], xi[1], obs_x[t], obs_y[t]);
            }
            VO_PHASE(0);
            double dz;
            ess = vo_weight_add_log(N, W, w, &dz);
            logz += dz;
            VO_PHASE(1);
        }
        /* do_resample: record ESS before resampling, then
         * Particle::resample (particle.hpp:170-197) */
        if (ess_out)
            ess_out[t] = ess;
        const int resampled = ess < cfg->threshold * (double)N;
        if (resampled) {
            vo_resample(cfg->scheme, N, N, &rng, W, rep);
            vo_trans_rep_index(N, N, rep, idx);
            VO_PHASE(2);
            vo_copy_rowmajor(N, 4, x, idx);
            VO_PHASE(3);
            ess = vo_weight_set_equal(N, W);
            VO_PHASE(1);
        }
        if (resampled_out)
            resampled_out[t] = resampled;
        /* monitor "pos" at MonitorMCMC: dgemv over (x, y) */
        if (pos_out) {
            for (size_t i = 0; i < N; ++i) {
                buf[2 * i] = x[4 * i];
                buf[2 * i + 1] = x[4 * i + 1];
            }
            vo_monitor_moments(N, 2, buf, W, pos_out + 2 * t);
            VO_PHASE(4);
        }
        if (logz_out)
            logz_out[t] = logz;
    }
#undef VO_PHASE
    const double elapsed = now_seconds() - t0;
    if (state_out)
        memcpy(state_out, x, sizeof(double) * 4 * N);
    if (W_out)
        memcpy(W_out, W, sizeof(double) * N);
    free(eng);
    free(x);
    free(W);
    free(w);
    free(buf);
    free(rep);
    free(idx);
    return elapsed;
}

/* ---------------------------------------------------------------- Kalman */

/* Dense 4-state Kalman filter for x_t = F x_{t-1} + q, q ~ N(0, Q),
 * y_t = H x_t + r, r ~ N(0, sd_obs^2 I2), x_0 ~ N(0, diag(sd_pos0^2,
 * sd_pos0^2, sd_vel0^2, sd_vel0^2)); the first observation is of x_0.
 * Textbook predict/update; log p(y_t | y_<t) from the 2x2 innovation. */
double vo_kalman_cv(int64_t T, double dt, double sd_pos, double sd_vel,
    double sd_pos0, double sd_vel0, double sd_obs, const double *obs_x,
    const double *obs_y, double *logz_out)
{
    double m[4] = {0, 0, 0, 0};
    double P[4][4] = {{0}};
    P[0][0] = P[1][1] = sd_pos0 * sd_pos0;
    P[2][2] = P[3][3] = sd_vel0 * sd_vel0;
    const double F[4][4] = {{1, 0, dt, 0}, {0, 1, 0, dt}, {0, 0, 1, 0},
        {0, 0, 0, 1}};
    const double Qd[4] = {sd_pos * sd_pos, sd_pos * sd_pos, sd_vel * sd_vel,
        sd_vel * sd_vel};
    const double R = sd_obs * sd_obs;
    double ll = 0;
    for (int64_t t = 0; t < T; ++t) {
        if (t > 0) {
            double mp[4] = {0, 0, 0, 0};
            double FP[4][4] = {{0}};
            double Pp[4][4] = {{0}};
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    mp[a] += F[a][b] * m[b];
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    for (int c = 0; c < 4; ++c)
                        FP[a][b] += F[a][c] * P[c][b];
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    for (int c = 0; c < 4; ++c)
                        Pp[a][b] += FP[a][c] * F[b][c];
                }
            for (int a = 0; a < 4; ++a)
                Pp[a][a] += Qd[a];
            memcpy(m, mp, sizeof(m));
            memcpy(P, Pp, sizeof(P));
        }
        /* innovation v = y - H m, S = H P H' + R (2x2) */
        const double v0 = obs_x[t] - m[0];
        const double v1 = obs_y[t] - m[1];
        const double S00 = P[0][0] + R, S01 = P[0][1], S10 = P[1][0],
                     S11 = P[1][1] + R;
        const double det = S00 * S11 - S01 * S10;
        const double i00 = S11 / det, i01 = -S01 / det, i10 = -S10 / det,
                     i11 = S00 / det;
        const double quad = v0 * (i00 * v0 + i01 * v1) +
            v1 * (i10 * v0 + i11 * v1);
        ll += -0.5 * quad - 0.5 * log(det) - log(two_pi);
        if (logz_out)
            logz_out[t] = ll;
        /* K = P H' S^-1 (4x2); m += K v; P -= K H P */
        double K[4][2];
        for (int a = 0; a < 4; ++a) {
            K[a][0] = P[a][0] * i00 + P[a][1] * i10;
            K[a][1] = P[a][0] * i01 + P[a][1] * i11;
        }
        for (int a = 0; a < 4; ++a)
            m[a] += K[a][0] * v0 + K[a][1] * v1;
        double Pn[4][4];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                Pn[a][b] = P[a][b] - (K[a][0] * P[0][b] + K[a][1] * P[1][b]);
        memcpy(P, Pn, sizeof(P));
    }
    return ll;
}

/* The sequential left fold s_i = (((w0 + w1) + w2) + ... + wi) that
 * resample_trans_u01_rep accumulates (include/vsmc/resample/transform.hpp:
 * 61-63). */
void vo_cumsum(size_t n, const double *w, double *s)
{
    double acc = 0;
    for (size_t i = 0; i < n; ++i) {
        acc += w[i];
        s[i] = acc;
    }
}
