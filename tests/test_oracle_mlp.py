"""Pins the learned-proposal oracle (oracle/mlp.c) without the device:
bfloat16 rounding against an independent integer formulation, the loss
gradient against central finite differences of the loss (FP64 mode), closed
forms at a zero last layer, and the proposal's importance-weight increment
against log densities from scipy."""
import math
import struct

import numpy as np
import pytest
from scipy import stats

import oracle_lib as ol


def _bf16_ref(x):
    """Round float32(x) to bfloat16, ties to even, by comparing the dropped
    16 bits with one half (independent of the oracle's add-and-mask)."""
    u = struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]
    hi, lo = u >> 16, u & 0xFFFF
    if lo > 0x8000 or (lo == 0x8000 and (hi & 1)):
        hi += 1
    return struct.unpack("<f", struct.pack("<I", (hi << 16) & 0xFFFFFFFF))[0]


def test_bf16_rounding():
    one = 1.0
    assert ol.bf16(one) == 1.0
    assert ol.bf16(1 + 2 ** -8) == 1.0                    # tie -> even (down)
    assert ol.bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6        # tie -> even (up)
    assert ol.bf16(1 + 2 ** -8 + 2 ** -20) == 1 + 2 ** -7  # above the tie
    assert ol.bf16(-2.5) == -2.5
    rs = np.random.default_rng(0)
    for x in np.concatenate([rs.normal(0, 3, 3000), rs.normal(0, 1e-20, 100),
                             [0.0, -0.0, 65504.0, 3.0e38]]):
        assert ol.bf16(x) == _bf16_ref(x), x
    assert math.isinf(ol.bf16(3.4e38))


def _rand_params(rs, scale=0.3):
    p = rs.normal(0, scale, ol.MLP_NPARAMS).astype(np.float32)
    return p


def _data(rs, n):
    feat = np.zeros((n, 8))
    feat[:, :6] = rs.normal(0, 1, (n, 6))
    target = rs.normal(0, 1, (n, 4))
    return feat, target


def test_gradient_finite_differences():
    rs = np.random.default_rng(1)
    p = _rand_params(rs)
    shift = rs.normal(0, 0.1, 6)
    scale = rs.uniform(0.5, 2, 6)
    feat, target = _data(rs, 5)
    loss, g = ol.mlp_loss_grad(p, shift, scale, feat, target, round_bf16=False)
    idx = np.concatenate([rs.choice(ol.MLP_NPARAMS, 150, replace=False),
                          np.arange(5120, 5128), np.arange(0, 12)])
    for i in idx:
        h = 1e-2 * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] = np.float32(p[i] + h)
        pm[i] = np.float32(p[i] - h)
        dp = float(pp[i]) - float(pm[i])
        lp, _ = ol.mlp_loss_grad(pp, shift, scale, feat, target, round_bf16=False)
        lm, _ = ol.mlp_loss_grad(pm, shift, scale, feat, target, round_bf16=False)
        fd = (lp - lm) / dp
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_zero_last_layer_closed_form():
    """W3 = b3 = 0: outputs 0 (the bootstrap proposal), loss = mean sum t^2/2,
    dL/db3 = (-mean t, mean(1 - t^2)), dL/dW3 = dL/db3 (x) mean h2 rows."""
    rs = np.random.default_rng(2)
    p = _rand_params(rs)
    o, _ = ol.MLP_OFF["W3"]
    p[o:o + 8 * 64 + 8] = 0
    feat, target = _data(rs, 9)
    out = ol.mlp_forward(p, np.zeros(6), np.ones(6), feat, round_bf16=False)
    assert np.all(out == 0)
    loss, g = ol.mlp_loss_grad(p, np.zeros(6), np.ones(6), feat, target, round_bf16=False)
    assert abs(loss - 0.5 * (target ** 2).sum(axis=1).mean()) < 1e-12
    gb3 = g[5120:5128]
    np.testing.assert_allclose(gb3[:4], -target.mean(axis=0), atol=1e-12)
    np.testing.assert_allclose(gb3[4:], (1 - target ** 2).mean(axis=0), atol=1e-12)
    # the hidden layers get no gradient through a zero last layer
    assert np.all(g[:4608] == 0)


def test_forward_rounding_close():
    """The bfloat16 mode stays within the rounding it models of FP64."""
    rs = np.random.default_rng(3)
    p = _rand_params(rs, 0.2)
    feat, _ = _data(rs, 50)
    a = ol.mlp_forward(p, np.zeros(6), np.ones(6), feat, round_bf16=False)
    b = ol.mlp_forward(p, np.zeros(6), np.ones(6), feat, round_bf16=True)
    assert np.abs(a - b).max() < 0.05 * np.abs(a).max()
    assert np.abs(a - b).max() > 0


def test_features():
    f = ol.cv_features(1.0, 2.0, 0.5, -1.0, 1.2, 1.7, dt=0.1)
    np.testing.assert_allclose(f, [1.2 - 1.05, 1.7 - 1.9, 0.5, -1.0, 0.2, -0.3], atol=1e-15)


@pytest.mark.parametrize("obs_model", [0, 1, 2])
def test_learned_particle_weight(obs_model):
    """x_t = m + s (mu + sigma z) and log w = log g + log f - log q, with f
    and q diagonal Gaussians evaluated independently by scipy."""
    rs = np.random.default_rng(4 + obs_model)
    dt, sp, sv = 0.1, 0.02, 0.03
    for _ in range(50):
        s0 = rs.normal(0, 1, 4)
        o = rs.normal(0, 0.7, 8)
        z = rs.normal(0, 1, 4)
        ox, oy = rs.normal(0, 1, 2)
        s1, inc = ol.cv_learned_particle(s0, ox, oy, o, z, dt, sp, sv, obs_model,
                                         0.1, 10.0)
        m = np.array([s0[0] + dt * s0[2], s0[1] + dt * s0[3], s0[2], s0[3]])
        sc = np.array([sp, sp, sv, sv])
        floor = float(np.float32(-0.3465)) if obs_model == 1 else -5.0
        ls = np.clip(o[4:], floor, 3)
        mu_q = m + sc * o[:4]
        sd_q = sc * np.exp(ls)
        np.testing.assert_allclose(s1, mu_q + sd_q * z, rtol=1e-14, atol=1e-14)
        logf = stats.norm.logpdf(s1, m, sc).sum()
        logq = stats.norm.logpdf(s1, mu_q, sd_q).sum()
        if obs_model == 2:
            logg = stats.norm.logpdf([ox, oy], s1[:2], 0.1).sum()
        else:
            logg = stats.t.logpdf([ox, oy], 10.0, s1[:2], 0.1).sum()
            if obs_model == 0:
                # the reference's form drops the constants
                logg -= 2 * (math.lgamma(5.5) - math.lgamma(5.0)
                             - 0.5 * math.log(10 * math.pi) - math.log(0.1))
        assert abs(inc - (logg + logf - logq)) < 1e-9, (inc, logg + logf - logq)


def test_heavy_tail_scale_floor():
    """Under the Student-t likelihood (obs_model 1) the proposal's log-scale
    is floored just above -ln(2)/2 (e^(2 ls) > 1/2: p g / q then has a finite
    second moment); the Gaussian likelihood keeps the -5 floor."""
    floor = float(np.float32(-0.3465))
    assert 0.5 < math.exp(2 * floor) < 0.5002
    s0 = np.array([0.3, -0.2, 1.0, 0.5])
    z = np.array([0.1, -1.2, 0.7, 2.0])
    o = np.concatenate([np.zeros(4), np.full(4, -3.0)])
    m = np.array([s0[0] + 0.1 * s0[2], s0[1] + 0.1 * s0[3], s0[2], s0[3]])
    sc = np.array([0.02, 0.02, 0.02, 0.02])
    for model, ls in ((1, floor), (2, -3.0), (0, -3.0)):
        s1, _ = ol.cv_learned_particle(s0, 0.4, -0.1, o, z, 0.1, 0.02, 0.02, model, 0.1, 10.0)
        np.testing.assert_allclose(s1, m + sc * np.exp(ls) * z, rtol=1e-14, atol=1e-15)


def test_bootstrap_proposal_weight():
    """o = 0: the proposal is the transition, the increment is log g."""
    s0 = np.array([0.3, -0.2, 1.0, 0.5])
    z = np.array([0.1, -1.2, 0.7, 2.0])
    s1, inc = ol.cv_learned_particle(s0, 0.4, -0.1, np.zeros(8), z)
    assert abs(inc - ol.cv_loglik(s1[0], s1[1], 0.4, -0.1)) < 1e-12
