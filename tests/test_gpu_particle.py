"""Device particle set (Weight, Particle::resample, Monitor) and the whole
constant-velocity PF loop against the oracle's Sampler restatement."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def _np(t):
    return t.cpu().numpy()


def c0_obs():
    with open(os.path.join(GOLD, "cv_obs_c0.json")) as f:
        g = json.load(f)
    return (np.array([float.fromhex(x) for x in g["obs_x"]]),
            np.array([float.fromhex(x) for x in g["obs_y"]]))


@pytest.mark.parametrize("n", [1, 3, 4096, 100003])
def test_weights(dev, n):
    import vsmc_amd as va
    rs = np.random.default_rng(n)
    p = va.Particle(n, 2, "float64", seed=5)
    lw = rs.normal(0, 3, n)
    p.set_log(_t(lw))
    W, ess, dz = ol.weight_set_log(lw)
    np.testing.assert_allclose(_np(p.weights()), W, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(p.ess(), ess, rtol=1e-12)
    np.testing.assert_allclose(p.log_zconst(), dz, rtol=1e-12, atol=1e-12)
    inc = rs.normal(0, 1, n)
    p.add_log(_t(inc))
    W2, ess2, dz2 = ol.weight_add_log(W, inc)
    np.testing.assert_allclose(_np(p.weights()), W2, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(p.ess(), ess2, rtol=1e-11)
    np.testing.assert_allclose(p.log_zconst(), dz + dz2, rtol=1e-12, atol=1e-12)
    p.set_equal()
    assert abs(p.ess() - n) < 1e-9 * n
    p.close()


@pytest.mark.parametrize("scheme", range(6))
def test_particle_resample(dev, scheme):
    """Particle::resample on the device: decision, counts from the device
    weights with the resampling engine (key seed_key(seed, N)), ancestors,
    in-place gather of the SoA state, equal weights afterwards."""
    import torch
    import vsmc_amd as va
    n = 50000
    rs = np.random.default_rng(scheme)
    p = va.Particle(n, 3, "float32", seed=11)
    x = torch.as_tensor(rs.normal(size=(3, n)), dtype=torch.float32, device="cuda")
    p.set_state(x)
    p.set_log(_t(rs.normal(0, 2, n)))
    W = _np(p.weights())
    _, rkey, rword = p.keys()
    assert rword == 0
    resampled, rep, idx = p.resample(scheme, 0.5, want_rep=True)
    assert resampled
    rep, idx = _np(rep), _np(idx)
    erep, used = ol.resample(scheme, W, n, rkey)
    if scheme in (ol.MULTINOMIAL, ol.RESIDUAL):
        assert int((rep != erep).sum()) <= 2
    else:
        assert (rep == erep).all()
    assert (idx == ol.trans_rep_index(rep)).all()
    assert (_np(p.state()) == _np(x)[:, idx]).all()
    assert abs(p.ess() - n) < 1e-6 * n
    _, _, rword = p.keys()
    if scheme not in (ol.MULTINOMIAL, ol.RESIDUAL):
        assert rword == used
    # threshold never: nothing happens
    r2, _, _ = p.resample(scheme, -math.inf)
    assert not r2
    p.close()


@pytest.mark.parametrize("s", [0.5, 2.0, 8.0])
def test_particle_events_full_size(dev, s):
    """Particle-set events at N = 2^24 with log-weights ~ N(0, s^2) for the
    schemes whose draws are sorted uniforms or residuals (the particle set
    counts them on its own passes): counts bit-exact with the standalone
    resampler fed the same device weights and stream position (which
    test_full_size_multinomial_residual pins to the oracle), words consumed
    equal, the ancestors the rep -> index layout; the residual stratified and
    systematic counts and words also bit-exact with the oracle."""
    import torch
    import vsmc_amd as va
    n = 1 << 24
    rs = np.random.default_rng(int(s * 1000) + 5)
    lw = _t(rs.normal(0, s, n))
    for scheme in (ol.MULTINOMIAL, ol.RESIDUAL, ol.RESIDUAL_STRATIFIED,
                   ol.RESIDUAL_SYSTEMATIC):
        print(f"scheme={scheme} s={s} n={n}", flush=True)
        # a new particle set per scheme: its stream starts at word 0
        p = va.Particle(n, 1, "float32", seed=5 + scheme)
        p.set_state(torch.zeros((1, n), dtype=torch.float32, device="cuda"))
        p.set_log(lw)
        Wd = p.weights()
        _, rkey, rword = p.keys()
        resampled, rep, idx = p.resample(scheme, math.inf, want_rep=True)
        assert resampled
        rep = _np(rep)
        erep, off2 = va.resample(scheme, Wd, n, rkey, rword)
        bad = np.nonzero(rep != _np(erep))[0]
        assert len(bad) == 0, (scheme, len(bad), bad[:5])
        assert rep.sum() == n
        _, _, rword2 = p.keys()
        assert rword2 == off2
        if scheme in (ol.RESIDUAL_STRATIFIED, ol.RESIDUAL_SYSTEMATIC):
            orep, used = ol.resample(scheme, _np(Wd), n, rkey, skip=rword)
            assert (rep == orep).all()
            assert rword2 - rword == used
        assert (_np(idx) == ol.trans_rep_index(rep)).all()
        p.close()


def test_weight_exponential(dev):
    """The weights the resampler uses, W_i = e^(l_i - lmax) / sum with the
    library's table-based exponential: against math.exp over [-700, 0]
    (every 64th of an octave, both signs of the remainder, integers and
    random points), W_i / W_max within 3 ulp of e^(l_i - lmax)."""
    import vsmc_amd as va
    rs = np.random.default_rng(7)
    ln2 = math.log(2.0)
    x = np.concatenate([
        -np.arange(0, 700, dtype=np.float64),
        -(np.arange(64 * 40) / 64.0) * ln2,
        -(np.arange(64 * 40) / 64.0 + 1.0 / 128.0) * ln2,
        -rs.uniform(0, 700, 20000),
        -rs.uniform(0, 1, 20000)])
    l = 3.25 + x                       # lmax = 3.25 (x = 0 is present)
    p = va.Particle(len(l), 1, "float64", seed=3)
    p.set_log(_t(l))
    W = _np(p.weights())
    imax = int(np.argmax(l))
    ratio = W / W[imax]
    ref = np.array([math.exp(v) for v in x])
    err = np.abs(ratio - ref) / (ref * 2.0 ** -52)
    assert err.max() <= 3.0, (float(err.max()), float(x[np.argmax(err)]))
    p.close()


def test_monitor_moments(dev):
    import torch
    import vsmc_amd as va
    n = 70001
    rs = np.random.default_rng(2)
    p = va.Particle(n, 4, "float64", seed=1)
    x = rs.normal(size=(4, n))
    p.set_state(torch.as_tensor(x, device="cuda"))
    p.set_log(_t(rs.normal(0, 1, n)))
    W = _np(p.weights())
    r = p.moments(1, 2)
    np.testing.assert_allclose(r, x[1:3] @ W, rtol=1e-11)


@pytest.mark.parametrize("scheme", [ol.STRATIFIED, ol.SYSTEMATIC, ol.MULTINOMIAL,
                                    ol.RESIDUAL])
def test_pf_c0_vs_oracle(dev, scheme):
    """CPU workload (N = 10^4, T = 100, FP64, ESS < N/2): the device run and
    the oracle agree step by step: resampled flags equal, ESS and the
    monitor "pos" within 1e-9 relative, log-evidence within 1e-9."""
    import vsmc_amd as va
    ox, oy = c0_obs()
    N, T, seed = 10000, 100, 20161
    f = va.CvFilter(N, "float64", seed=seed)
    g = f.run(ox, oy, scheme, 0.5)
    e = ol.pf_run(N, T, seed, scheme, 0.5, ox, oy)
    print("resampled steps", int(g["resampled"].sum()), int(e["resampled"].sum()), flush=True)
    assert (g["resampled"] == e["resampled"]).all()
    np.testing.assert_allclose(g["ess"], e["ess"], rtol=1e-9)
    np.testing.assert_allclose(g["pos"], e["pos"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(g["logz"], e["logz"], rtol=1e-9, atol=1e-9)
    f.close()


def test_pf_fp32_large(dev):
    """FP32 state, FP64 weights at N = 2^20: finite, tracks the FP64 run's
    log-evidence within Monte Carlo error."""
    import vsmc_amd as va
    ox, oy = c0_obs()
    f = va.CvFilter(1 << 20, "float32", seed=3)
    g = f.run(ox[:50], oy[:50], ol.SYSTEMATIC, 0.5)
    f.close()
    h = va.CvFilter(1 << 20, "float64", seed=3)
    k = h.run(ox[:50], oy[:50], ol.SYSTEMATIC, 0.5)
    h.close()
    assert np.isfinite(g["logz"]).all()
    assert abs(g["logz"][-1] - k["logz"][-1]) < 2.0


def _cv_normals_f32(words):
    """The FP32 path's Box-Muller pairs in FP64 from the same words: FP32
    uniforms (u >> 8) 2^-24 (u01.hpp:126-147), radius sqrt(-2 log ua),
    angle 2 pi ub."""
    u = (words >> 8).astype(np.float64) * 2.0 ** -24
    r0, a0 = np.sqrt(-2 * np.log(u[0])), 2 * np.pi * u[1]
    r1, a1 = np.sqrt(-2 * np.log(u[2])), 2 * np.pi * u[3]
    return r0 * np.cos(a0), r0 * np.sin(a0), r1 * np.cos(a1), r1 * np.sin(a1)


def _student_ll(x, y, ox, oy, scale=0.1, nu=10.0):
    rx, ry = (x - ox) / scale, (y - oy) / scale
    return -0.5 * (nu + 1) * (np.log1p(rx * rx / nu) + np.log1p(ry * ry / nu))


@pytest.mark.parametrize("n", [4096, 10002])
def test_cv_fp32_step_math(dev, n):
    """FP32 state: init and one move against the model equations
    (pf_seq.cpp:70-81, :111-122) evaluated in FP64 from the same Philox
    words: state within FP32 rounding of the arithmetic (2e-6), log-weights
    within 5e-5. n = 10002 also runs the scalar tail path."""
    import vsmc_amd as va
    ox, oy = c0_obs()
    f = va.CvFilter(n, "float32", seed=9)
    k0 = f.keys()[0]
    keys = (k0 - 1 + np.arange(n, dtype=np.int64)) % 0xFFFFFFFF + 1
    f.init(float(ox[0]), float(oy[0]))
    s0 = _np(f.state()).astype(np.float64)
    W0 = _np(f.weights())
    f.move(1, float(ox[1]), float(oy[1]))
    s1 = _np(f.state()).astype(np.float64)
    W1 = _np(f.weights())
    f.close()
    w = np.array([ol.philox_words(int(k), 0, 8) for k in keys], dtype=np.uint64).T
    z = _cv_normals_f32(w[0:4])
    e0 = np.stack([2.0 * z[0], 2.0 * z[1], 1.0 * z[2], 1.0 * z[3]])
    np.testing.assert_allclose(s0, e0, rtol=2e-6, atol=2e-6)
    z = _cv_normals_f32(w[4:8])
    e1 = np.stack([s0[0] + (0.02 * z[0] + 0.1 * s0[2]), s0[1] + (0.02 * z[1] + 0.1 * s0[3]),
                   s0[2] + 0.02 * z[2], s0[3] + 0.02 * z[3]])
    np.testing.assert_allclose(s1, e1, rtol=2e-6, atol=2e-6)
    l0 = _student_ll(s0[0], s0[1], ox[0], oy[0])
    l1 = l0 + _student_ll(s1[0], s1[1], ox[1], oy[1])
    for W, l in ((W0, l0), (W1, l1)):
        lw = l - l.max()
        lw -= np.log(np.exp(lw).sum())
        keep = lw > -600
        np.testing.assert_allclose(np.log(W[keep]), lw[keep], rtol=0, atol=5e-5)
