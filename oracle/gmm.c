/* gmm.c -- CPU oracle (test infrastructure only) of the tempered SMC
 * sampler for a Gaussian mixture: example/gmm/include/gmm.hpp of the
 * reference, restated per particle with the random draws passed in, so
 * that the device's moves can be checked draw for draw.
 *
 * Parameters of one particle (c components): mu[c], lambda[c] (precisions),
 * weight[c] (simplex). Prior (gmm.hpp:302-314): mu ~ N(mu0, sd0^2), lambda
 * with log density (shape0 - 1) log lambda - lambda / scale0, weights flat
 * on the simplex; constants dropped as in the reference.
 * Log-likelihood (gmm.hpp:316-331):
 *   -n/2 log(2 pi) + sum_k log(sum_i w_i exp(log(lambda_i)/2
 *                                         - lambda_i (x_k - mu_i)^2 / 2) + 1e-13).
 * Hyper-parameters from the data range (gmm.hpp:333-357): kappa =
 * 1 / range^2, mu0 = mid-range, sd0 = 1 / sqrt(kappa), shape0 = 2,
 * scale0 = 50 kappa.
 * Moves (gmm.hpp:447-543), each with target log p = log prior + alpha log
 * likelihood: mu += sd_mu z; lambda *= exp(sd_lambda z) with Jacobian
 * sum log(lambda' / lambda); logit-weight random walk (log(w_i / w_c) +=
 * sd_w z for i < c) with Jacobian logit_weight_diff (gmm.hpp:122-143).
 * Accept when log u >= p (mh_reject_common, gmm.hpp:176-185).
 * Proposal scales (gmm.hpp:424-428) from alpha clamped below at 0.02:
 * sd_mu = 0.15 / alpha, sd_lambda = (1 + sqrt(1 / alpha)) 0.15,
 * sd_w = (1 + sqrt(1 / alpha)) 0.2.
 *
 * Adaptive tempering (no counterpart in the reference, which anneals on a
 * fixed schedule, gmm.hpp:567-604): the conditional ESS of an increment d,
 *   CESS(d) = (sum_i W_i e^{d l_i})^2 / sum_i W_i e^{2 d l_i},
 * and the largest d in (0, 1 - alpha] with CESS(d) >= target.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LN_2PI 1.8378770664093454835606594728112352797227949472755668

void vo_gmm_hyper(const double *x, int64_t n, double h[4])
{
    double lo = x[0], hi = x[0];
    for (int64_t k = 0; k < n; ++k) {
        lo = x[k] < lo ? x[k] : lo;
        hi = x[k] > hi ? x[k] : hi;
    }
    const double kappa = 1 / ((hi - lo) * (hi - lo));
    h[0] = 0.5 * (lo + hi);   /* mu0 */
    h[1] = 1 / sqrt(kappa);   /* sd0 */
    h[2] = 2;                 /* shape0 */
    h[3] = 50 * kappa;        /* scale0 */
}

double vo_gmm_log_prior(int c, const double *mu, const double *lambda,
    const double h[4])
{
    double lp = 0;
    for (int i = 0; i < c; ++i) {
        const double r = mu[i] - h[0];
        lp += -0.5 * (r * r) / (h[1] * h[1]);
        lp += (h[2] - 1) * log(lambda[i]) - lambda[i] / h[3];
    }
    return lp;
}

double vo_gmm_log_lik(int c, const double *mu, const double *lambda,
    const double *weight, const double *x, int64_t n)
{
    double ll = -0.5 * (double)n * LN_2PI;
    for (int64_t k = 0; k < n; ++k) {
        double s = 0;
        for (int i = 0; i < c; ++i) {
            const double r = x[k] - mu[i];
            s += weight[i] * exp(0.5 * log(lambda[i]) - 0.5 * lambda[i] * r * r);
        }
        ll += log(s + 1e-13);
    }
    return ll;
}

void vo_gmm_scales(double alpha, double sd[3])
{
    const double a = alpha < 0.02 ? 0.02 : alpha;
    sd[0] = 0.15 / a;
    sd[1] = (1 + sqrt(1 / a)) * 0.15;
    sd[2] = (1 + sqrt(1 / a)) * 0.2;
}

/* log-logit weight sum of gmm_param::logit_weight_diff (one side). */
static double logit_side(int c, const double *w)
{
    double sum_lw = 1, sum_llw = 0;
    for (int i = 0; i < c - 1; ++i) {
        const double r = w[i] / w[c - 1];
        sum_lw += r;
        sum_llw += log(r);
    }
    return sum_llw - c * log(sum_lw);
}

/* One MH move of kind 0 (mu), 1 (lambda) or 2 (weight) on one particle:
 * p[3c + 2] = (mu, lambda, weight, log prior, log likelihood), z[c]
 * standard normals (c - 1 used for weights), log_u = log of the uniform.
 * Returns 1 when accepted (p updated), 0 when rejected (p unchanged). */
int vo_gmm_move(int kind, int c, double *p, const double *z, double log_u,
    double alpha, const double h[4], const double *x, int64_t n)
{
    double q[3 * 8 + 2];
    memcpy(q, p, sizeof(double) * (3 * c + 2));
    double *mu = q, *lambda = q + c, *w = q + 2 * c;
    double sd[3];
    vo_gmm_scales(alpha, sd);
    const double p0 = p[3 * c] + alpha * p[3 * c + 1];
    double jac = 0;
    if (kind == 0) {
        for (int i = 0; i < c; ++i)
            mu[i] += sd[0] * z[i];
    } else if (kind == 1) {
        for (int i = 0; i < c; ++i) {
            const double f = exp(sd[1] * z[i]);
            lambda[i] *= f;
            jac += log(lambda[i] / p[c + i]);
        }
    } else {
        double sum = 1;
        for (int i = 0; i < c - 1; ++i) {
            w[i] = exp(log(w[i] / w[c - 1]) + sd[2] * z[i]);
            sum += w[i];
        }
        w[c - 1] = 1;
        for (int i = 0; i < c; ++i)
            w[i] /= sum;
        jac = logit_side(c, w) - logit_side(c, p + 2 * c);
    }
    q[3 * c] = vo_gmm_log_prior(c, mu, lambda, h);
    q[3 * c + 1] = vo_gmm_log_lik(c, mu, lambda, w, x, n);
    const double pr = jac + q[3 * c] + alpha * q[3 * c + 1] - p0;
    if (pr < log_u)
        return 0;
    memcpy(p, q, sizeof(double) * (3 * c + 2));
    return 1;
}

double vo_cess(int64_t n, const double *W, const double *ll, double d)
{
    double m = -INFINITY;
    for (int64_t i = 0; i < n; ++i)
        m = ll[i] > m ? ll[i] : m;
    double s1 = 0, s2 = 0;
    for (int64_t i = 0; i < n; ++i) {
        /* e^{0 ll} = 1, also for ll = -inf */
        const double e = d == 0 ? 1.0 : exp(d * (ll[i] - m));
        s1 += W[i] * e;
        s2 += W[i] * e * e;
    }
    return s1 * s1 / s2;
}

/* Largest d in (0, dmax] with CESS(d) >= target, by bisection to tol. */
double vo_cess_solve(int64_t n, const double *W, const double *ll,
    double dmax, double target, double tol)
{
    if (vo_cess(n, W, ll, dmax) >= target)
        return dmax;
    double lo = 0, hi = dmax;
    while (hi - lo > tol) {
        const double mid = 0.5 * (lo + hi);
        if (vo_cess(n, W, ll, mid) >= target)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* vo_gmm_move over n particles (rows of 3c + 2 in p, c normals per particle
 * in z, log-uniforms in lu) on `threads` OpenMP threads (0: all); returns
 * the number accepted. The CPU baseline of the device's moves. */
int64_t vo_gmm_moves(int kind, int c, int64_t n, double *p, const double *z,
    const double *lu, double alpha, const double h[4], const double *x,
    int64_t nd, int threads)
{
    int64_t acc = 0;
#ifdef _OPENMP
    if (threads > 0)
        omp_set_num_threads(threads);
#pragma omp parallel for reduction(+ : acc) schedule(static)
#endif
    for (int64_t i = 0; i < n; ++i)
        acc += vo_gmm_move(kind, c, p + i * (3 * c + 2), z + i * c, lu[i], alpha,
            h, x, nd);
    return acc;
}
