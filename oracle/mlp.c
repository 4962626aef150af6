/* mlp.c -- CPU oracle (test infrastructure only) of the learned Gaussian
 * proposal for the constant-velocity filter: a 6 -> 64 -> 64 -> 8 tanh MLP
 * mapping (previous state, observation) to the proposal's mean offsets and
 * log-scale offsets, its Gaussian negative log-likelihood loss and the
 * loss gradient, and the importance-weight increment p g / q of one
 * proposed particle.
 *
 * The reference has no learned proposal; the model it corrects is the PF of
 * user_guide/cpp/pf_seq.cpp:70-81, :111-122 (transition) and
 * example/pf/include/pf.hpp:78-87 (observation density), and the weight
 * update is Weight::add_log (include/vsmc/core/weight.hpp:200-229).
 *
 * Two modes: round = 0 evaluates everything in FP64 (pinned in the tests by
 * central finite differences of the loss and by closed forms); round = 1
 * additionally rounds, to bfloat16 with round-to-nearest-even, exactly the
 * values the device's matrix-core products consume (inputs, weights,
 * hidden activations, and in the backward pass the output and
 * pre-activation gradients), so the device can be compared with it.
 *
 * Parameter vector (float, kParams = 5128):
 *   W1[64][6] | b1[64] | W2[64][64] | b2[64] | W3[8][64] | b3[8]
 * Features f[6] of (x, y, vx, vy) and (ox, oy) with dt:
 *   ox - (x + dt vx), oy - (y + dt vy), vx, vy, ox - x, oy - y,
 * normalised as (f - shift) * scale before the first layer. Outputs
 * o[0..3] are the mean offsets in units of the prior scales and o[4..7] the
 * log-scale offsets: u_d = o_d + exp(o_{4+d}) z_d, x_t = m + s * u with
 * the prior mean m = (x + dt vx, y + dt vy, vx, vy) and scale
 * s = (sd_pos, sd_pos, sd_vel, sd_vel).
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#define H 64
#define NI 6
#define NO 8
#define OW1 0
#define OB1 (OW1 + H * NI)
#define OW2 (OB1 + H)
#define OB2 (OW2 + H * H)
#define OW3 (OB2 + H)
#define OB3 (OW3 + NO * H)
#define NPAR (OB3 + NO)

int vo_mlp_nparams(void) { return NPAR; }

/* float -> bfloat16 -> float, round to nearest even (NaN kept quiet). */
double vo_bf16(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) {
        u = (u | 0x00400000u) & 0xFFFF0000u;
    } else {
        u += 0x7FFFu + ((u >> 16) & 1u);
        u &= 0xFFFF0000u;
    }
    memcpy(&f, &u, 4);
    return (double)f;
}

static double rb(double x, int round) { return round ? vo_bf16(x) : x; }

void vo_cv_features(double x, double y, double vx, double vy, double ox,
    double oy, double dt, double f[6])
{
    f[0] = ox - (x + dt * vx);
    f[1] = oy - (y + dt * vy);
    f[2] = vx;
    f[3] = vy;
    f[4] = ox - x;
    f[5] = oy - y;
}

struct fwd {
    double in[NI];
    double h1[H], h2[H];   /* unrounded tanh outputs */
    double o[NO];
};

static void forward1(const float *p, const double *shift, const double *scale,
    const double *feat, int round, struct fwd *r)
{
    for (int k = 0; k < NI; ++k)
        r->in[k] = rb((feat[k] - shift[k]) * scale[k], round);
    for (int i = 0; i < H; ++i) {
        double a = p[OB1 + i];
        for (int k = 0; k < NI; ++k)
            a += rb(p[OW1 + i * NI + k], round) * r->in[k];
        r->h1[i] = tanh(a);
    }
    for (int i = 0; i < H; ++i) {
        double a = p[OB2 + i];
        for (int k = 0; k < H; ++k)
            a += rb(p[OW2 + i * H + k], round) * rb(r->h1[k], round);
        r->h2[i] = tanh(a);
    }
    for (int i = 0; i < NO; ++i) {
        double a = p[OB3 + i];
        for (int k = 0; k < H; ++k)
            a += rb(p[OW3 + i * H + k], round) * rb(r->h2[k], round);
        r->o[i] = a;
    }
}

/* out[n][8]; feat[n][8] (6 used). */
void vo_mlp_forward(const float *params, const double *shift,
    const double *scale, int64_t n, const double *feat, int round, double *out)
{
    struct fwd r;
    for (int64_t s = 0; s < n; ++s) {
        forward1(params, shift, scale, feat + 8 * s, round, &r);
        memcpy(out + 8 * s, r.o, sizeof(r.o));
    }
}

/* Mean over n samples of sum_d [ls_d + (r_d)^2 / 2], r_d = (t_d - mu_d)
 * exp(-ls_d) (the Gaussian negative log-likelihood without 2 pi), and its
 * gradient with respect to the parameter vector (grad[kParams]). Returns
 * the mean loss. */
double vo_mlp_loss_grad(const float *params, const double *shift,
    const double *scale, int64_t n, const double *feat, const double *target,
    int round, double *grad)
{
    memset(grad, 0, sizeof(double) * NPAR);
    double loss = 0;
    struct fwd r;
    for (int64_t s = 0; s < n; ++s) {
        forward1(params, shift, scale, feat + 8 * s, round, &r);
        double go[NO];
        for (int d = 0; d < 4; ++d) {
            const double mu = r.o[d], ls = r.o[4 + d];
            const double e = exp(-ls);
            const double rr = (target[4 * s + d] - mu) * e;
            loss += ls + 0.5 * rr * rr;
            go[d] = -rr * e / (double)n;
            go[4 + d] = (1 - rr * rr) / (double)n;
        }
        double gob[NO];
        for (int i = 0; i < NO; ++i)
            gob[i] = rb(go[i], round);
        /* layer 3 */
        for (int i = 0; i < NO; ++i) {
            grad[OB3 + i] += go[i];
            for (int k = 0; k < H; ++k)
                grad[OW3 + i * H + k] += gob[i] * rb(r.h2[k], round);
        }
        double gz2[H];
        for (int k = 0; k < H; ++k) {
            double g = 0;
            for (int i = 0; i < NO; ++i)
                g += rb(params[OW3 + i * H + k], round) * gob[i];
            gz2[k] = g * (1 - r.h2[k] * r.h2[k]);
        }
        double gz2b[H];
        for (int k = 0; k < H; ++k)
            gz2b[k] = rb(gz2[k], round);
        for (int i = 0; i < H; ++i) {
            grad[OB2 + i] += gz2[i];
            for (int k = 0; k < H; ++k)
                grad[OW2 + i * H + k] += gz2b[i] * rb(r.h1[k], round);
        }
        double gz1[H];
        for (int k = 0; k < H; ++k) {
            double g = 0;
            for (int i = 0; i < H; ++i)
                g += rb(params[OW2 + i * H + k], round) * gz2b[i];
            gz1[k] = g * (1 - r.h1[k] * r.h1[k]);
        }
        for (int i = 0; i < H; ++i) {
            grad[OB1 + i] += gz1[i];
            const double g = rb(gz1[i], round);
            for (int k = 0; k < NI; ++k)
                grad[OW1 + i * NI + k] += g * r.in[k];
        }
    }
    return loss / (double)n;
}

/* One proposed particle: from the previous state s0[4], the observation,
 * the network outputs o[8] and the standard normals z[4], the new state
 * s1[4] and the log-weight increment
 *   log g(y | x_t) + log f(x_t | x_{t-1}) - log q(x_t | x_{t-1}, y)
 * = ll(x_t) + sum_d [ls_d + z_d^2 / 2 - u_d^2 / 2],
 * with ls_d clamped to [-5, 3] as on the device, the lower end raised to
 * -0.3465 (just above -ln(2) / 2, so e^(2 ls) > 1/2) for the Student-t
 * likelihood with constants (obs_model 1): below it p g / q has an infinite
 * second moment under polynomial tails. obs_model / scale / df as
 * vo_pf_config (0: Student-t without constants, 1: with, 2: Gaussian). */
double vo_cv_learned_particle(const double s0[4], double ox, double oy,
    const double o[8], const double z[4], double dt, double sd_pos,
    double sd_vel, int obs_model, double obs_scale, double obs_df,
    double s1[4])
{
    const double m[4] = {s0[0] + dt * s0[2], s0[1] + dt * s0[3], s0[2], s0[3]};
    const double sc[4] = {sd_pos, sd_pos, sd_vel, sd_vel};
    double inc = 0;
    const double ls_min = obs_model == 1 ? (double)-0.3465f : -5;
    for (int d = 0; d < 4; ++d) {
        double ls = o[4 + d];
        ls = ls < ls_min ? ls_min : (ls > 3 ? 3 : ls);
        const double u = o[d] + exp(ls) * z[d];
        s1[d] = m[d] + sc[d] * u;
        inc += ls + 0.5 * z[d] * z[d] - 0.5 * u * u;
    }
    return inc + vo_cv_loglik(s1[0], s1[1], ox, oy, obs_model, obs_scale, obs_df);
}
