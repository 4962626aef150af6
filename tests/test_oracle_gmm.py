"""Pins the tempered mixture SMC oracle (oracle/gmm.c) on the CPU: the
log-likelihood and log prior against scipy densities (with the constants
the reference drops), the hyper-parameters from the data range, the
Jacobians of the multiplicative-precision and logit-weight moves against
closed forms, the accept rule, and the conditional ESS and its solver
against their definitions and a brute-force search."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle_lib as ol


def _data(n=200, seed=0):
    rs = np.random.default_rng(seed)
    comp = rs.choice(4, n, p=[0.1, 0.2, 0.3, 0.4])
    mu = np.array([-3.0, 0.0, 2.0, 5.0])
    sd = np.array([0.5, 1.5, 0.75, 1.0])
    return rs.normal(mu[comp], sd[comp])


def _params(rs, c=4):
    mu = rs.normal(0, 3, c)
    lam = rs.gamma(2, 0.5, c)
    w = rs.dirichlet(np.ones(c))
    return mu, lam, w


def test_hyper():
    x = _data()
    h = ol.gmm_hyper(x)
    rng = x.max() - x.min()
    np.testing.assert_allclose(h, [0.5 * (x.max() + x.min()), rng, 2.0, 50 / rng ** 2],
                               rtol=1e-14)


def test_log_lik_is_mixture_density():
    x = _data()
    rs = np.random.default_rng(1)
    for _ in range(20):
        mu, lam, w = _params(rs)
        dens = (w[None, :] * stats.norm.pdf(x[:, None], mu[None, :],
                                            1 / np.sqrt(lam[None, :]))).sum(axis=1)
        # + 1e-13 inside the log, times sqrt(2 pi) from the dropped constant
        exp = np.log(dens * math.sqrt(2 * math.pi) + 1e-13).sum() - 0.5 * len(x) * \
            math.log(2 * math.pi)
        assert abs(ol.gmm_log_lik(mu, lam, w, x) - exp) < 1e-9 * abs(exp)


def test_log_prior_is_normal_gamma():
    x = _data()
    h = ol.gmm_hyper(x)
    rs = np.random.default_rng(2)
    for _ in range(20):
        mu, lam, _ = _params(rs)
        exp = (stats.norm.logpdf(mu, h[0], h[1]) + 0.5 * math.log(2 * math.pi * h[1] ** 2)
               + stats.gamma.logpdf(lam, h[2], scale=h[3]) + math.lgamma(h[2])
               + h[2] * math.log(h[3])).sum()
        assert abs(ol.gmm_log_prior(mu, lam, h) - exp) < 1e-9 * max(1, abs(exp))


def test_scales():
    np.testing.assert_allclose(ol.gmm_scales(0.5), [0.3, (1 + math.sqrt(2)) * 0.15,
                                                     (1 + math.sqrt(2)) * 0.2])
    np.testing.assert_allclose(ol.gmm_scales(0.001), ol.gmm_scales(0.02))


def _state(mu, lam, w, h, x):
    return np.concatenate([mu, lam, w, [ol.gmm_log_prior(mu, lam, h),
                                        ol.gmm_log_lik(mu, lam, w, x)]])


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_move_accept_rule_and_jacobian(kind):
    """The move's log acceptance ratio equals
    Jacobian + (log prior + alpha log lik)' - (log prior + alpha log lik),
    with the Jacobian sum log(lambda'/lambda) (precisions) or
    sum_i log w'_i - sum_i log w_i (logit weights: softmax Jacobian);
    a log-uniform just below the ratio accepts, just above rejects."""
    x = _data(100, 3)
    h = ol.gmm_hyper(x)
    rs = np.random.default_rng(4 + kind)
    alpha = 0.37
    sd = ol.gmm_scales(alpha)
    for _ in range(10):
        mu, lam, w = _params(rs)
        p = _state(mu, lam, w, h, x)
        z = rs.normal(0, 1, 4)
        mu2, lam2, w2 = mu.copy(), lam.copy(), w.copy()
        if kind == 0:
            mu2 = mu + sd[0] * z
            jac = 0.0
        elif kind == 1:
            lam2 = lam * np.exp(sd[1] * z)
            jac = np.log(lam2 / lam).sum()
        else:
            eta = np.log(w[:3] / w[3]) + sd[2] * z[:3]
            r = np.concatenate([np.exp(eta), [1.0]])
            w2 = r / r.sum()
            jac = np.log(w2).sum() - np.log(w).sum()
        q = _state(mu2, lam2, w2, h, x)
        ratio = jac + q[12] + alpha * q[13] - p[12] - alpha * p[13]
        lu_acc = min(ratio, 0.0) - 1e-9
        acc, got = ol.gmm_move(kind, p, z, lu_acc, alpha, h, x)
        assert acc == 1
        np.testing.assert_allclose(got, q, rtol=1e-12, atol=1e-12)
        if ratio < -1e-6:
            acc, got = ol.gmm_move(kind, p, z, ratio + 1e-7, alpha, h, x)
            assert acc == 0 and np.array_equal(got, p)


def test_zero_step_always_accepts():
    x = _data(50, 5)
    h = ol.gmm_hyper(x)
    mu, lam, w = _params(np.random.default_rng(6))
    p = _state(mu, lam, w, h, x)
    for kind in range(3):
        acc, got = ol.gmm_move(kind, p, np.zeros(4), -1e-300, 0.5, h, x)
        assert acc == 1
        np.testing.assert_allclose(got, p, rtol=1e-13)


def test_cess():
    rs = np.random.default_rng(7)
    n = 1000
    W = rs.random(n)
    W /= W.sum()
    ll = rs.normal(-500, 20, n)
    assert abs(ol.cess(W, np.full(n, 3.0), 0.7) - 1.0) < 1e-14
    for d in (0.0, 1e-3, 0.05, 0.3):
        e = np.exp(d * (ll - ll.max()))
        exp = (W @ e) ** 2 / (W @ (e * e))
        assert abs(ol.cess(W, ll, d) - exp) < 1e-12
    # solver: largest d with CESS >= 0.9, against a fine grid
    d = ol.cess_solve(W, ll, 1.0, 0.9)
    grid = np.linspace(0, 0.05, 50001)
    ok = grid[[ol.cess(W, ll, g) >= 0.9 for g in grid]]
    assert abs(d - ok.max()) <= 0.05 / 50000 + 1e-12
    assert ol.cess(W, ll, d) >= 0.9 - 1e-12
    assert ol.cess_solve(W, np.full(n, 1.0), 0.25, 0.9) == 0.25
