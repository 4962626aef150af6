"""Pins the random-walk Metropolis-Hastings oracle (oracle/rw.c), which
restates include/vsmc/rng/random_walk.hpp:49-555.

The reference's own header cannot be compiled here (it includes
include/vsmc/math/cblas.h, which needs a system CBLAS or MKL), so the oracle
is pinned by what the mathematics fixes:
  * flat targets accept every proposal, so the chain is the start plus the
    stream's normals: rebuilt here from raw Philox words with Box-Muller
    (normal_distribution.hpp:199-216) and the cached second half;
  * the bounded transforms are log / logit shifts by z, and their log
    q-ratio is the Jacobian ratio, checked by finite differences;
  * long chains reach the targets' known moments (a dropped Jacobian term
    or a wrong sign moves them by many standard errors);
  * each particle's word count follows from the draw order.
"""
import math

import numpy as np
import pytest

import oracle_lib as o

M32 = 0xFFFFFFFF
TWO_PI = 6.283185307179586


def seed_key(s, i):
    return ((s % M32) + i) % M32 + 1


class Stream:
    """Words of one particle's engine, read in order, with the cached
    Box-Muller half of a NormalDistribution(0, sd)."""

    def __init__(self, key):
        self.w = o.philox_words(key, 0, 4096)
        self.pos = 0
        self.saved = None

    def word(self):
        v = int(self.w[self.pos])
        self.pos += 1
        return v

    def normal(self, sd=1.0):
        if self.saved is not None:
            z, self.saved = self.saved, None
        else:
            u1 = self.word() / 4294967296.0
            u2 = self.word() / 4294967296.0
            r = math.sqrt(-2 * math.log(u1))
            z = r * math.cos(TWO_PI * u2)
            self.saved = r * math.sin(TWO_PI * u2)
        return 0.0 + sd * z


CHOL3 = np.array([[0.5, 0, 0], [0.2, 0.4, 0], [0.1, -0.1, 0.3]])


def test_flat_chain_is_the_stream():
    seed, n, K = 77, 6, 9
    out = o.rw_run(o.RW_FLAT, seed, n, 1, K)
    for i in range(n):
        st = Stream(seed_key(seed, i))
        x = 0.0
        for _ in range(K):
            x = x + st.normal(1.3)
            st.word()  # the uniform of the accept test
        assert out[i, 4] == K
        assert out[i, 7] == st.pos
        assert out[i, 0] == pytest.approx(x, rel=0, abs=1e-13)


def test_flat_mv_chain_is_chol_times_stream():
    seed, n, K = 5, 4, 7
    out = o.rw_run(o.RW_FLAT_MV, seed, n, 1, K)
    for i in range(n):
        st = Stream(seed_key(seed, i))
        x = np.array([1.0, 2.0, 3.0])
        for _ in range(K):
            z = np.array([st.normal() for _ in range(3)])
            x = x + CHOL3 @ z
            st.word()
        assert out[i, 4] == K
        assert out[i, 7] == st.pos
        np.testing.assert_allclose(out[i, :3], x, rtol=0, atol=1e-12)


def test_flags():
    inf = math.inf
    assert o.rw_flag(-inf, inf) == 0
    assert o.rw_flag(-inf, 0.0) == 1
    assert o.rw_flag(0.0, inf) == 2
    assert o.rw_flag(0.0, 1.0) == 3


def _logit(v, a, b):
    return math.log((v - a) / (b - v))


def test_transform_closed_forms():
    rng = np.random.default_rng(3)
    for _ in range(200):
        z = float(rng.normal(0, 1.5))
        # unbounded: shift by z, ratio 1
        x = float(rng.normal(0, 3))
        y, q = o.rw_q(0, x, z)
        assert y == x + z and q == 0
        # lower bound a: log(y - a) = log(x - a) + z
        a = float(rng.normal(0, 2))
        x = a + float(rng.exponential(2.0))
        y, q = o.rw_q(2, x, z, a, math.inf)
        assert y > a or y == a  # underflow of exp(z) only
        assert math.log(y - a) == pytest.approx(math.log(x - a) + z, abs=1e-12)
        assert q == z
        # upper bound b: log(b - y) = log(b - x) + z
        b = float(rng.normal(0, 2))
        x = b - float(rng.exponential(2.0))
        y, q = o.rw_q(1, x, z, -math.inf, b)
        assert math.log(b - y) == pytest.approx(math.log(b - x) + z, abs=1e-12)
        assert q == z
        # both: logit(y) = logit(x) + z
        a = float(rng.normal(0, 2))
        b = a + float(rng.exponential(3.0)) + 0.1
        x = a + (b - a) * float(rng.uniform(0.05, 0.95))
        zz = float(np.clip(z, -4, 4))
        y, q = o.rw_q(3, x, zz, a, b)
        assert a < y < b
        assert _logit(y, a, b) == pytest.approx(_logit(x, a, b) + zz, abs=1e-9)


@pytest.mark.parametrize("flag", [1, 2, 3])
def test_q_ratio_is_the_jacobian(flag):
    """The proposal is a Gaussian random walk in eta(x) (log or logit), so
    q(y | x) = N(eta(y) - eta(x)) |eta'(y)| and the ratio the accept test
    needs, q(x | y) / q(y | x), is |eta'(x)| / |eta'(y)|. Derivatives by
    central differences of eta."""
    a, b = -1.0, 2.5
    eta = {1: lambda v: math.log(b - v), 2: lambda v: math.log(v - a),
           3: lambda v: _logit(v, a, b)}[flag]
    lo = a if flag in (2, 3) else -math.inf
    hi = b if flag in (1, 3) else math.inf
    rng = np.random.default_rng(flag)
    for _ in range(50):
        x = -0.5 + 2.5 * float(rng.uniform(0.02, 0.98))
        if flag == 2:
            x = a + float(rng.exponential(1.0))
        if flag == 1:
            x = b - float(rng.exponential(1.0))
        z = float(rng.normal(0, 0.8))
        y, q = o.rw_q(flag, x, z, lo, hi)

        def d(v):
            h = 1e-6 * max(1.0, abs(v))
            h = min(h, 0.25 * min(abs(v - a), abs(b - v)))
            return abs(eta(v + h) - eta(v - h)) / (2 * h)
        # q(x|y)/q(y|x) = |eta'(x)| / |eta'(y)|
        assert q == pytest.approx(math.log(d(x) / d(y)), abs=1e-6)


def _words_expected(scenario, iters, steps):
    """Each move call builds its proposal afresh: ceil(dim steps / 2)
    Box-Muller pairs plus one uniform per step."""
    dim = {o.RW_MV3: 3, o.RW_BLOCK: 2, o.RW_FLAT_MV: 3}.get(scenario, 1)
    return iters * (2 * ((dim * steps + 1) // 2) + steps)


@pytest.mark.parametrize("scenario", range(o.RW_NSCEN))
def test_word_counts(scenario):
    iters, steps = 3, 5
    out = o.rw_run(scenario, 11, 64, iters, steps)
    assert (out[:, 7] == _words_expected(scenario, iters, steps)).all()
    assert (out[:, 4] >= 0).all() and (out[:, 4] <= iters * steps).all()


def _lt(scenario, x):
    if scenario in (o.RW_NORMAL, o.RW_G):
        return -0.5 * x[0] * x[0]
    if scenario == o.RW_GAMMA:
        return 1.5 * math.log(x[0]) - x[0]
    if scenario == o.RW_BETA:
        return math.log(x[0]) + 2 * math.log(1 - x[0])
    if scenario == o.RW_NEG_GAMMA:
        return 1.5 * math.log(-x[0]) + x[0]
    if scenario in (o.RW_MV3, o.RW_BLOCK, o.RW_BLOCK1):
        d = x[1] - 0.5 * x[0]
        return -0.5 * (x[0] ** 2 + d * d / 0.5) + math.log(x[2]) - x[2]
    return 0.0


@pytest.mark.parametrize("scenario", range(o.RW_NSCEN))
def test_log_target_carried(scenario):
    """ltx holds log gamma of the current state after any accept/reject
    sequence (updated on accept only)."""
    out = o.rw_run(scenario, 23, 200, 2, 7)
    dim = o.oracle().vo_rw_dim(scenario)
    for r in out[:50]:
        assert r[3] == pytest.approx(_lt(scenario, r[:dim]), abs=1e-12)


def test_g_is_the_test_function_of_the_state():
    out = o.rw_run(o.RW_G, 9, 500, 1, 20)
    assert (out[:, 5] == out[:, 0]).all()
    assert (out[:, 6] == out[:, 0] * out[:, 0]).all()


def test_block_leaves_other_components():
    n = 300
    out = o.rw_run(o.RW_BLOCK, 4, n, 2, 10)
    out1 = o.rw_run(o.RW_BLOCK1, 4, n, 2, 10)
    for i in range(n):
        x0 = o.rw_init(o.RW_BLOCK, i)
        assert out[i, 0] == x0[0]
        assert out1[i, 0] == x0[0] and out1[i, 1] == x0[1]


def _z(est, mean, sd_mean):
    return abs(est - mean) / sd_mean


@pytest.mark.parametrize("scenario,mean,var", [
    (o.RW_NORMAL, 0.0, 1.0),
    (o.RW_G, 0.0, 1.0),
    (o.RW_GAMMA, 2.5, 2.5),
    (o.RW_BETA, 0.4, 0.04),
    (o.RW_NEG_GAMMA, -2.5, 2.5),
])
def test_stationary_moments_1d(scenario, mean, var):
    n = 20000
    out = o.rw_run(scenario, 1234, n, 1, 300)
    x = out[:, 0]
    assert _z(x.mean(), mean, math.sqrt(var / n)) < 5
    # variance of the sample variance ~ (m4 - var^2) / n; 4 var^2 / n bounds
    # it for these targets
    assert _z(x.var(), var, math.sqrt(4 * var * var / n)) < 5


def test_stationary_moments_mv():
    n = 20000
    out = o.rw_run(o.RW_MV3, 99, n, 1, 300)
    x = out[:, :3]
    sd = lambda v: math.sqrt(v / n)
    assert _z(x[:, 0].mean(), 0, sd(1)) < 5
    assert _z(x[:, 1].mean(), 0, sd(0.75)) < 5
    assert _z(x[:, 2].mean(), 2, sd(2)) < 5
    assert _z(x[:, 0].var(), 1, sd(4)) < 5
    assert _z(x[:, 1].var(), 0.75, sd(4 * 0.5625)) < 5
    assert _z(x[:, 2].var(), 2, sd(4 * 4 * 2)) < 5
    c01 = np.mean(x[:, 0] * x[:, 1]) - x[:, 0].mean() * x[:, 1].mean()
    assert _z(c01, 0.5, sd(2)) < 5
    # block updates: x1 | x0 ~ N(x0 / 2, 1/2), x2 ~ Gamma(2)
    out = o.rw_run(o.RW_BLOCK, 98, n, 1, 300)
    r = out[:, 1] - 0.5 * out[:, 0]
    assert _z(r.mean(), 0, sd(0.5)) < 5
    assert _z(r.var(), 0.5, sd(4 * 0.25)) < 5
    assert _z(out[:, 2].mean(), 2, sd(2)) < 5
    out = o.rw_run(o.RW_BLOCK1, 97, n, 1, 300)
    assert _z(out[:, 2].mean(), 2, sd(2)) < 5
    assert _z(out[:, 2].var(), 2, sd(4 * 4 * 2)) < 5


def test_unknown_scenario():
    with pytest.raises(ValueError):
        o.rw_run(42, 1, 4, 1, 1)
