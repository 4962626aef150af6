"""Tempered SMC for a Gaussian mixture on the device against the oracle
(oracle/gmm.c): initial log prior / log likelihood of every particle, each
Metropolis-Hastings move draw for draw (same Philox words), the first
tempering increment against the conditional-ESS solver, and an end-to-end
run (alpha reaches 1, the two evidence estimates agree, the posterior
recovers the mixture)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _data(n=200, seed=5):
    from vsmc_amd.models import gmm_simulate
    return gmm_simulate(n, seed)


def _keys(s, n):
    k0 = s.particle.keys()[0]
    return (k0 - 1 + np.arange(n, dtype=np.int64)) % 0xFFFFFFFF + 1


def _normals_f32(words):
    """FP32 Box-Muller pairs of consecutive words, recomputed in FP64."""
    u = (np.asarray(words, dtype=np.uint64) >> 8).astype(np.float64) * 2.0 ** -24
    z = []
    for a, b in zip(u[0::2], u[1::2]):
        r = np.sqrt(-2 * np.log(a))
        z += [r * np.cos(2 * np.pi * b), r * np.sin(2 * np.pi * b)]
    return np.array(z)


def test_init_values(dev):
    import vsmc_amd as va
    x = _data()
    h = ol.gmm_hyper(x)
    s = va.GmmSampler(4096, x, ncomp=4, seed=3)
    r = s.run(max_steps=1)
    assert r["steps"] == 1
    st = s.state().cpu().numpy().astype(np.float64)
    W = s.weights().cpu().numpy()
    assert np.allclose(W, 1 / 4096)
    w = st[8:12]
    np.testing.assert_allclose(w.sum(axis=0), 1.0, rtol=1e-5)
    assert (st[4:8] > 0).all()
    for i in range(0, 4096, 97):
        lp = ol.gmm_log_prior(st[0:4, i], st[4:8, i], h)
        ll = ol.gmm_log_lik(st[0:4, i], st[4:8, i], st[8:12, i], x)
        assert abs(st[12, i] - lp) <= 1e-4 * max(1.0, abs(lp)), (i, st[12, i], lp)
        assert abs(st[13, i] - ll) <= 2e-5 * max(1.0, abs(ll)), (i, st[13, i], ll)
    # prior moments of the draws: mu ~ N(mu0, sd0), lambda ~ Gamma(2, scale0)
    assert abs(st[0:4].mean() - h[0]) < 0.05 * h[1]
    assert abs(st[4:8].mean() / (2 * h[3]) - 1) < 0.05
    s.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_move_vs_oracle(dev, kind):
    """One move of each kind at alpha = 0.3 from the initial particles, with
    the draws of step 1, move index kind: the accept decisions agree (up
    to FP32 rounding at the boundary) and accepted states match."""
    import vsmc_amd as va
    x = _data()
    h = ol.gmm_hyper(x)
    n = 2048
    s = va.GmmSampler(n, x, ncomp=4, seed=11)
    s.run(max_steps=1)
    keys = _keys(s, n)
    st0 = s.state().cpu().numpy().astype(np.float64)
    s.move(kind, 0.3, 1, kind)
    st1 = s.state().cpu().numpy().astype(np.float64)
    s.close()
    off = 64 * (32 * 1 + kind + 1)
    agree = 0
    for i in range(n):
        words = ol.philox_words(int(keys[i]), off, 6)
        z = _normals_f32(words[:4])
        u = (int(words[4]) >> 8) * 2.0 ** -24
        lu = np.log(u) if u > 0 else -np.inf
        acc, q = ol.gmm_move(kind, st0[:, i], z, lu, 0.3, h, x)
        moved = not np.array_equal(st1[:, i], st0[:, i])
        if bool(acc) == moved:
            agree += 1
            if acc:
                np.testing.assert_allclose(st1[:12, i], q[:12], rtol=2e-4, atol=1e-5)
                assert abs(st1[13, i] - q[13]) <= 1e-4 * max(1.0, abs(q[13]))
    assert agree >= n - 3, agree


@pytest.mark.parametrize("n", [1 << 14, (1 << 20) + 77])
def test_first_increment(dev, n):
    """From equal weights at alpha = 0, the device's increment equals the
    oracle's conditional-ESS solution on the same log-likelihoods (at 2^20 + 77
    the search's 512 blocks loop over several particles per thread, with a
    ragged end)."""
    import vsmc_amd as va
    x = _data()
    s = va.GmmSampler(n, x, ncomp=4, seed=9, cess=0.9)
    s.run(max_steps=1)
    ll = s.state().cpu().numpy()[13].astype(np.float64)
    r = s.run(max_steps=2)
    s.close()
    d = ol.cess_solve(np.full(n, 1.0 / n), ll, 1.0, 0.9, tol=1e-14)
    assert abs(r["alpha"][1] - d) <= 1e-6 * d + 1e-12, (r["alpha"][1], d)
    # the ESS after the reweight (before resampling) is the target
    assert abs(r["ess"][1] / n - 0.9) < 2e-3


def test_run_end_to_end(dev):
    """1,000 points, N = 2^14: alpha increases to 1; the SMC log-evidence
    and the path-sampling estimate agree; the posterior (components sorted
    by mean) recovers the mixture's means."""
    import vsmc_amd as va
    from vsmc_amd.models import gmm_simulate
    x = gmm_simulate(1000, 2016)
    s = va.GmmSampler(1 << 14, x, ncomp=4, seed=101, cess=0.9, moves=5)
    r = s.run(max_steps=400)
    st = s.state().cpu().numpy().astype(np.float64)
    W = s.weights().cpu().numpy()
    s.close()
    a = r["alpha"]
    print("steps", r["steps"], "logz", r["logz"][-1], "path", r["path_sampling"],
          "accept", r["accept"][-1], flush=True)
    assert a[-1] == 1.0
    assert (np.diff(a) > 0).all()
    assert np.isfinite(r["logz"]).all()
    assert abs(r["logz"][-1] - r["path_sampling"]) < 2.0
    acc = r["accept"][1:]
    assert ((acc > 0) & (acc < 1)).all()
    mu = np.sort(st[0:4], axis=0)
    post = (mu * W[None, :]).sum(axis=1)
    np.testing.assert_allclose(post, [-3.0, 0.0, 2.0, 5.0], atol=0.6)


def test_large_n_finite(dev):
    """N = 2^20 draws some FP32 uniforms of exactly 0 (infinite Box-Muller
    normals): their proposals are rejected, the state stays finite and the
    run reaches alpha = 1 in a few dozen steps."""
    import vsmc_amd as va
    from vsmc_amd.models import gmm_simulate
    x = gmm_simulate(1000, 2016)
    s = va.GmmSampler(1 << 20, x, ncomp=4, seed=101, cess=0.9, moves=5)
    r = s.run(max_steps=200)
    st = s.state().cpu().numpy()
    s.close()
    print("steps", r["steps"], "logz", r["logz"][-1], "path", r["path_sampling"], flush=True)
    assert r["alpha"][-1] == 1.0
    assert np.isfinite(st).all()
    assert np.isfinite(r["integrand"]).all()
    assert abs(r["logz"][-1] - r["path_sampling"]) < 2.0
