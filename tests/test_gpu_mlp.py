"""The learned proposal on the device: network forward and loss gradient of
the bfloat16 matrix-core kernels against the oracle (oracle/mlp.c in its
bfloat16 mode), training on transitions the bootstrap filter simulates,
one learned move against the oracle's per-particle formula, and the
trained filter's log-evidence error against the FP64 Kalman filter on the
linear-Gaussian model, next to the bootstrap filter's."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

# linear-Gaussian variant: CV dynamics, Gaussian position observations with
# sd 0.01, initial state known to the process-noise scale
LG = dict(dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=0.02, sd_vel0=0.02, obs_model=2,
          obs_scale=0.01, obs_df=10.0)


def _rand_net(rs, scale=0.3):
    p = rs.normal(0, scale, ol.MLP_NPARAMS).astype(np.float32)
    shift = rs.normal(0, 0.1, 6).astype(np.float32)
    sc = rs.uniform(0.5, 2.0, 6).astype(np.float32)
    return p, shift, sc


def _feat(rs, n):
    f = np.zeros((n, 8), dtype=np.float32)
    f[:, :6] = rs.normal(0, 1, (n, 6))
    return f


@pytest.mark.parametrize("n", [1, 15, 16, 1000, 70001])
def test_forward_vs_oracle(dev, n):
    import torch
    import vsmc_amd as va
    print(f"n={n}", flush=True)
    rs = np.random.default_rng(n)
    p, shift, sc = _rand_net(rs)
    net = va.ProposalNet()
    net.set_params(p, shift, sc)
    assert np.array_equal(net.params(), p)
    f = _feat(rs, n)
    got = net.forward(torch.as_tensor(f)).cpu().numpy()
    exp = ol.mlp_forward(p, shift.astype(np.float64), sc.astype(np.float64),
                         f.astype(np.float64), round_bf16=True)
    err = np.abs(got - exp)
    scale = np.abs(exp).max()
    # bfloat16 activations: a rounding that FP32 and FP64 accumulation
    # decide differently moves an output by about one bf16 ulp of a term
    assert err.max() <= 2e-2 * scale, (err.max(), scale)
    assert np.median(err) <= 1e-3 * scale, np.median(err)
    net.close()


def test_zero_net_is_bootstrap(dev):
    import torch
    import vsmc_amd as va
    net = va.ProposalNet()
    out = net.forward(torch.as_tensor(_feat(np.random.default_rng(0), 100))).cpu().numpy()
    assert np.all(out == 0)
    net.close()


@pytest.mark.parametrize("n", [64, 5000])
def test_loss_grad_vs_oracle(dev, n):
    import torch
    import vsmc_amd as va
    print(f"n={n}", flush=True)
    rs = np.random.default_rng(7 + n)
    p, shift, sc = _rand_net(rs, 0.2)
    net = va.ProposalNet()
    net.set_params(p, shift, sc)
    f = _feat(rs, n)
    t = rs.normal(0, 1, (n, 4)).astype(np.float32)
    loss, g = net.loss_grad(torch.as_tensor(f), torch.as_tensor(t))
    eloss, eg = ol.mlp_loss_grad(p, shift.astype(np.float64), sc.astype(np.float64),
                                 f.astype(np.float64), t.astype(np.float64), round_bf16=True)
    assert abs(loss - eloss) <= 1e-3 * abs(eloss), (loss, eloss)
    rel = np.linalg.norm(g - eg) / np.linalg.norm(eg)
    assert rel < 2e-2, rel
    # per block of parameters
    for name, (o, shape) in ol.MLP_OFF.items():
        k = int(np.prod(shape))
        a, b = g[o:o + k], eg[o:o + k]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-7, name
    net.close()


def _lg_track(T, seed):
    from vsmc_amd.models import cv_simulate
    _, ox, oy = cv_simulate(T, seed, dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=0.02,
                            sd_vel0=0.02, obs="gauss", obs_scale=0.01)
    return ox, oy


def _trained_lg_net(steps=600):
    import vsmc_amd as va
    ox, oy = _lg_track(64, 11)
    f = va.CvFilter(1 << 14, "float32", seed=5, **LG)
    feat, target = f.transitions(ox, oy, va.SYSTEMATIC, 0.5, train_seed=77)
    f.close()
    net = va.ProposalNet()
    net.fit_normalisation(feat)
    net.init_random(1)
    loss = net.train(feat, target, batch=4096, steps=steps, lr=2e-3)
    return net, loss, feat, target


def test_train_lg(dev):
    """Training on 63 x 2^14 transitions: the loss falls from the bootstrap
    proposal's E[sum z^2 / 2] = 2 toward the optimal proposal's
    2 (log(0.447) + 1/2) + 1 = 0.39 (position posterior / prior sd
    sqrt(0.01^2 / (0.02^2 + 0.01^2)))."""
    net, loss, feat, target = _trained_lg_net()
    print("loss", loss[:3], loss[-3:], flush=True)
    assert abs(loss[0] - 2.0) < 0.1
    assert np.isfinite(loss).all()
    assert loss[-20:].mean() < 0.6
    net.close()


# Student-t observations (df 10, scale 0.1) with 5% outliers at 10x scale,
# normalising constants kept (obs_model 1)
ST = dict(dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=0.02, sd_vel0=0.02, obs_model=1,
          obs_scale=0.1, obs_df=10.0)


def _st_track(T, seed):
    from vsmc_amd.models import cv_simulate
    _, ox, oy = cv_simulate(T, seed, dt=0.1, sd_pos=0.02, sd_vel=0.02, sd_pos0=0.02,
                            sd_vel0=0.02, obs="student", obs_scale=0.1, obs_df=10.0,
                            outlier_frac=0.05, outlier_mult=10.0)
    return ox, oy


@pytest.mark.parametrize("model", ["lg", "student"])
def test_learned_move_vs_oracle(dev, model):
    """One learned move: new states and log-weights against the oracle's
    per-particle formula with the device network's outputs and the
    particles' own normals (Philox block t + 1), for the linear-Gaussian and
    the Student-t (outlier) observation models."""
    import torch
    import vsmc_amd as va
    rs = np.random.default_rng(3)
    p, shift, sc = _rand_net(rs, 0.15)
    net = va.ProposalNet()
    net.set_params(p, shift, sc)
    n = 5000
    cfg = LG if model == "lg" else ST
    ox, oy = _lg_track(4, 2) if model == "lg" else _st_track(4, 2)
    f = va.CvFilter(n, "float32", seed=21, **cfg)
    k0 = f.keys()[0]
    keys = (k0 - 1 + np.arange(n, dtype=np.int64)) % 0xFFFFFFFF + 1
    f.init(float(ox[0]), float(oy[0]))
    s0 = f.state().cpu().numpy().astype(np.float64)
    W0 = f.weights().cpu().numpy()
    f.move_learned(net, 1, float(ox[1]), float(oy[1]))
    s1 = f.state().cpu().numpy().astype(np.float64)
    W1 = f.weights().cpu().numpy()
    f.close()
    feats = np.zeros((n, 8), dtype=np.float32)
    x, y, vx, vy = (s0[k].astype(np.float32) for k in range(4))
    o32, p32 = np.float32(ox[1]), np.float32(oy[1])
    dt = np.float32(0.1)
    feats[:, 0] = o32 - (x + dt * vx)
    feats[:, 1] = p32 - (y + dt * vy)
    feats[:, 2] = vx
    feats[:, 3] = vy
    feats[:, 4] = o32 - x
    feats[:, 5] = p32 - y
    out = net.forward(torch.as_tensor(feats)).cpu().numpy().astype(np.float64)
    w = np.array([ol.philox_words(int(k), 4, 4) for k in keys], dtype=np.uint64).T
    u = (w >> 8).astype(np.float64) * 2.0 ** -24
    r0, a0 = np.sqrt(-2 * np.log(u[0])), 2 * np.pi * u[1]
    r1, a1 = np.sqrt(-2 * np.log(u[2])), 2 * np.pi * u[3]
    z = np.stack([r0 * np.cos(a0), r0 * np.sin(a0), r1 * np.cos(a1), r1 * np.sin(a1)])
    es1 = np.zeros((4, n))
    inc = np.zeros(n)
    for i in range(n):
        es1[:, i], inc[i] = ol.cv_learned_particle(s0[:, i], float(o32), float(p32), out[i],
                                                   z[:, i], 0.1, 0.02, 0.02, cfg["obs_model"],
                                                   cfg["obs_scale"], cfg["obs_df"])
    np.testing.assert_allclose(s1, es1, rtol=1e-5, atol=2e-6)
    lw = np.log(W0) + inc
    lw -= lw.max()
    lw -= np.log(np.exp(lw).sum())
    keep = lw > -200
    np.testing.assert_allclose(np.log(W1[keep]), lw[keep], rtol=0, atol=2e-2)
    net.close()


def _rmse(net, n, seeds, T=50):
    import vsmc_amd as va
    errs = []
    for s in seeds:
        ox, oy = _lg_track(T, 1000 + s)
        kz, _ = ol.kalman_cv(ox, oy, 0.1, 0.02, 0.02, 0.02, 0.02, 0.01)
        f = va.CvFilter(n, "float32", seed=s, **LG)
        r = f.run(ox, oy, va.SYSTEMATIC, 0.5) if net is None else \
            f.run_learned(net, ox, oy, va.SYSTEMATIC, 0.5)
        f.close()
        errs.append(r["logz"][-1] - kz)
    return float(np.sqrt(np.mean(np.square(errs)))), errs


def test_learned_pf_beats_bootstrap_lg(dev):
    """Linear-Gaussian model, T = 50, N = 2^12, 6 tracks: the learned
    filter's log-evidence error against the FP64 Kalman filter is well
    below the bootstrap filter's at the same N (both unbiased in Z)."""
    net, loss, _, _ = _trained_lg_net()
    seeds = list(range(6))
    rb, eb = _rmse(None, 1 << 12, seeds)
    rl, el = _rmse(net, 1 << 12, seeds)
    print("bootstrap rmse", rb, eb, "learned rmse", rl, el, flush=True)
    assert np.isfinite(el).all()
    assert rl < 0.5 * rb
    assert rl < 1.0
    net.close()


def test_learned_pf_large_n(dev):
    """N = 2^19 over 50 steps draws ~2^25 FP32 uniforms: some are exactly 0
    (infinite Box-Muller radius). Such a particle gets weight 0 instead of
    a NaN log-weight, and the log-evidence stays close to the Kalman
    value; the monitor stays finite."""
    import vsmc_amd as va
    net, _, _, _ = _trained_lg_net(steps=300)
    ox, oy = _lg_track(50, 1003)
    kz, _ = ol.kalman_cv(ox, oy, 0.1, 0.02, 0.02, 0.02, 0.02, 0.01)
    for learned in (False, True):
        f = va.CvFilter(1 << 19, "float32", seed=8, **LG)
        r = f.run_learned(net, ox, oy, va.SYSTEMATIC, 0.5) if learned else \
            f.run(ox, oy, va.SYSTEMATIC, 0.5)
        W = f.weights().cpu().numpy()
        f.close()
        print("learned" if learned else "bootstrap", r["logz"][-1] - kz, flush=True)
        assert np.isfinite(W).all()
        assert np.isfinite(r["pos"]).all()
        assert abs(r["logz"][-1] - kz) < 0.5
    net.close()


def test_collapse_falls_back_to_bootstrap(dev):
    """A network whose proposal means sit 50 standard deviations off makes a
    learned step's ESS collapse (below N / 1000). From then on the filter's
    learned moves take the bootstrap proposal: from the same state, a
    learned move equals the bootstrap move (to FP32 rounding of the two
    orders of operations), while before any collapse it does not."""
    import vsmc_amd as va
    n = 1 << 14
    p = np.zeros(ol.MLP_NPARAMS, np.float32)
    p[ol.MLP_NPARAMS - 8:ol.MLP_NPARAMS - 4] = 50.0   # b3: the mean offsets
    net = va.ProposalNet()
    net.set_params(p, np.zeros(6, np.float32), np.ones(6, np.float32))
    ox, oy = _lg_track(4, 7)

    def start(collapse):
        f = va.CvFilter(n, "float32", seed=3, **LG)
        f.init(float(ox[0]), float(oy[0]))
        f.resample(va.SYSTEMATIC, 0.5)
        if collapse:
            f.move_learned(net, 1, float(ox[1]), float(oy[1]))
            assert f.ess() < n / 1000
            f.resample(va.SYSTEMATIC, 0.5)
        return f

    a, b = start(True), start(True)
    a.move(2, float(ox[2]), float(oy[2]))                 # bootstrap kernel
    b.move_learned(net, 2, float(ox[2]), float(oy[2]))    # after the collapse
    sa, sb = a.state().cpu().numpy(), b.state().cpu().numpy()
    wa, wb = a.weights().cpu().numpy(), b.weights().cpu().numpy()
    np.testing.assert_allclose(sb, sa, rtol=1e-5, atol=1e-6)
    keep = wa > 1e-200
    # log-likelihoods near -90 (1.0 from an observation with sd 0.01): the
    # FP32 state's rounding moves them by up to about 1e-4 relative
    np.testing.assert_allclose(np.log(wb[keep]), np.log(wa[keep]), rtol=1e-4, atol=1e-3)
    a.close()
    b.close()
    # without a collapse the learned move is the network's
    c, d = start(False), start(False)
    c.move(1, float(ox[1]), float(oy[1]))
    d.move_learned(net, 1, float(ox[1]), float(oy[1]))
    assert np.abs(d.state().cpu().numpy() - c.state().cpu().numpy()).max() > 0.5
    c.close()
    d.close()
    net.close()
