"""Weights, the PF loop and the Kalman filter of the oracle, pinned by closed
forms, invariants and a brute-force Gaussian marginal likelihood."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
from vsmc_amd.models import cv_simulate

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def c0_obs():
    with open(os.path.join(GOLD, "cv_obs_c0.json")) as f:
        g = json.load(f)
    return (np.array([float.fromhex(x) for x in g["obs_x"]]),
            np.array([float.fromhex(x) for x in g["obs_y"]]))


def test_c0_obs_fixture_reproducible():
    ox, oy = c0_obs()
    _, sx, sy = cv_simulate(100, 20161)
    assert (sx == ox).all() and (sy == oy).all()


def test_weight_set_log_closed_form():
    rs = np.random.default_rng(0)
    for n in (1, 2, 7, 1000, 10000):
        lw = rs.normal(0, 3, n)
        W, ess, dz = ol.weight_set_log(lw)
        p = np.exp(lw - lw.max())
        p /= p.sum()
        np.testing.assert_allclose(W, p, rtol=1e-12)
        assert abs(W.sum() - 1) < 1e-12
        assert 1 - 1e-9 <= ess <= n + 1e-9
        np.testing.assert_allclose(ess, 1 / (p ** 2).sum(), rtol=1e-12)
        # set_log after equal weights: log mean exp
        np.testing.assert_allclose(dz, np.log(np.mean(np.exp(lw - lw.max()))) + lw.max(),
                                   rtol=1e-12, atol=1e-12)


def test_weight_add_log_closed_form():
    rs = np.random.default_rng(1)
    n = 500
    W0, _, _ = ol.weight_set_log(rs.normal(0, 1, n))
    inc = rs.normal(0, 2, n)
    W, ess, dz = ol.weight_add_log(W0, inc)
    p = W0 * np.exp(inc)
    np.testing.assert_allclose(dz, np.log(p.sum()), rtol=1e-12)
    np.testing.assert_allclose(W, p / p.sum(), rtol=1e-12)
    np.testing.assert_allclose(ess, 1 / ((p / p.sum()) ** 2).sum(), rtol=1e-12)
    # equal weights
    We = np.zeros(n)
    assert abs(ol.oracle().vo_weight_set_equal(n, We) - n) < 1e-9 * n
    assert np.allclose(We, 1 / n)


def test_pf_first_step_closed_form():
    """T = 1, never resample: logZ_0 = log mean_i exp(ll_i) of the initial
    particles, and the state is 4 Box-Muller normals of block 1."""
    ox, oy = c0_obs()
    N = 2000
    out = ol.pf_run(N, 1, 20161, ol.STRATIFIED, -math.inf, ox, oy, want_state=True)
    x = out["state"]
    lx = np.log(1 + (10 * (x[:, 0] - ox[0])) ** 2 / 10)
    ly = np.log(1 + (10 * (x[:, 1] - oy[0])) ** 2 / 10)
    ll = -5.5 * (lx + ly)
    m = ll.max()
    np.testing.assert_allclose(out["logz"][0], m + np.log(np.mean(np.exp(ll - m))),
                               rtol=1e-12)
    # particle 0: key seed + 1, words 0..3 -> x = 2 r cos, y = 2 r sin
    u = ol.u01_fill(0, 20162, 0, 4)
    r1 = np.sqrt(-2 * np.log(u[0]))
    np.testing.assert_allclose(x[0, 0], 2 * r1 * np.cos(2 * np.pi * u[1]), rtol=1e-14)
    np.testing.assert_allclose(x[0, 1], 2 * r1 * np.sin(2 * np.pi * u[1]), rtol=1e-14)
    np.testing.assert_allclose(out["pos"][0], out["W"] @ x[:, :2], rtol=1e-12)


def test_pf_c0_runs_and_tracks():
    ox, oy = c0_obs()
    st, _, _ = cv_simulate(100, 20161)
    out = ol.pf_run(10000, 100, 20161, ol.STRATIFIED, 0.5, ox, oy)
    assert np.isfinite(out["logz"]).all()
    assert (out["ess"] > 0).all() and (out["ess"] <= 10000 * (1 + 1e-9)).all()
    assert out["resampled"].sum() > 0
    # the filter follows the simulated track once N is large enough for the
    # small process noise (at 10^4 a mid-run degeneracy episode is normal)
    out = ol.pf_run(100000, 100, 20161, ol.STRATIFIED, 0.5, ox, oy)
    err = np.abs(out["pos"] - st[:, :2]).mean()
    assert err < 0.2, err


def dense_gauss_loglik(ox, oy, dt, sp, sv, sp0, sv0, so):
    """log N(y; 0, H Sx H' + R) for the stacked state of T steps."""
    T = len(ox)
    F = np.array([[1, 0, dt, 0], [0, 1, 0, dt], [0, 0, 1, 0], [0, 0, 0, 1.0]])
    Q = np.diag([sp ** 2, sp ** 2, sv ** 2, sv ** 2])
    P0 = np.diag([sp0 ** 2, sp0 ** 2, sv0 ** 2, sv0 ** 2])
    # x_t = F^t x0 + sum_{k=1..t} F^{t-k} q_k
    A = [np.linalg.matrix_power(F, t) for t in range(T)]
    S = np.zeros((4 * T, 4 * T))
    for a in range(T):
        for b in range(T):
            c = A[a] @ P0 @ A[b].T
            for k in range(1, min(a, b) + 1):
                c += np.linalg.matrix_power(F, a - k) @ Q @ np.linalg.matrix_power(F, b - k).T
            S[4 * a:4 * a + 4, 4 * b:4 * b + 4] = c
    H = np.zeros((2 * T, 4 * T))
    for t in range(T):
        H[2 * t, 4 * t] = 1
        H[2 * t + 1, 4 * t + 1] = 1
    Sy = H @ S @ H.T + so ** 2 * np.eye(2 * T)
    y = np.ravel(np.column_stack([ox, oy]))
    sign, logdet = np.linalg.slogdet(Sy)
    return -0.5 * (y @ np.linalg.solve(Sy, y)) - 0.5 * logdet - T * math.log(2 * math.pi)


@pytest.mark.parametrize("T", [1, 2, 3])
def test_kalman_vs_dense(T):
    _, ox, oy = cv_simulate(T, 5, obs="gauss", obs_scale=0.01)
    ll, cum = ol.kalman_cv(ox, oy, sd_obs=0.01)
    exp = dense_gauss_loglik(ox, oy, 0.1, 0.02, 0.02, 2.0, 1.0, 0.01)
    np.testing.assert_allclose(ll, exp, rtol=1e-9)
    assert cum[-1] == ll


def test_pf_logz_vs_kalman():
    """Linear-Gaussian variant: the bootstrap PF's log-evidence estimates the
    Kalman value (Monte Carlo tolerance 1.5 nats at N = 20000, T = 20)."""
    _, ox, oy = cv_simulate(20, 11, obs="gauss", obs_scale=0.1)
    kf, _ = ol.kalman_cv(ox, oy, sd_obs=0.1)
    out = ol.pf_run(20000, 20, 99, ol.SYSTEMATIC, 0.5, ox, oy, obs_model=2)
    assert abs(out["logz"][-1] - kf) < 1.5, (out["logz"][-1], kf)
