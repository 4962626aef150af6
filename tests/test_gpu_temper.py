"""Adaptive tempering on the device (vsmc_amd/csrc/temper.hip,
vsmch_particle_cess / vsmch_particle_cess_solve) against the oracle's
conditional ESS and bisection (oracle/gmm.c vo_cess, vo_cess_solve; pinned
there by the definition on random weights, CESS = 1 for a constant
log-likelihood, and a 50,001-point grid for the solver). The reference
anneals on a fixed schedule (example/gmm/include/gmm.hpp:567-604), so there
is no reference output to match; tolerances: CESS to 1e-12 + 4 n 2^-53
relative (the oracle's left-to-right FP64 sums against the device's tree
sums), the increment to the search's resolution dmax / 16^rounds."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _setup(va, torch, n, seed, equal=False, ninf=0):
    rs = np.random.default_rng(seed)
    p = va.Particle(n, 1, "float64", seed=seed)
    if equal:
        p.set_equal()
    else:
        p.set_log(torch.as_tensor(rs.normal(0, 1.5, n), device="cuda"))
    ll = rs.normal(-3, 4, n)
    if ninf:
        ll[rs.choice(n, ninf, replace=False)] = -np.inf
    W = p.weights().cpu().numpy()
    return p, torch.as_tensor(ll, device="cuda"), ll, W


@pytest.mark.parametrize("n,equal,ninf", [(1, False, 0), (4095, False, 0), (100003, False, 7),
                                          (100003, True, 0), (1 << 20, False, 100)])
def test_cess_vs_oracle(dev, n, equal, ninf):
    import torch
    import vsmc_amd as va
    p, llt, ll, W = _setup(va, torch, n, 11 + n, equal, ninf)
    for d in (0.0, 1e-3, 0.05, 0.4, 1.0, 3.0):
        got = p.cess(llt, d)
        ref = ol.cess(W, ll, d)
        print(f"n={n} equal={equal} d={d} cess={got} oracle={ref}", flush=True)
        # the oracle sums left to right: its rounding error grows with n
        assert got == pytest.approx(ref, rel=1e-12 + 4 * n * 2.0 ** -53, abs=1e-300)
    # d = 0 is 1 whenever some particle has weight
    assert p.cess(llt, 0.0) == pytest.approx(1.0, rel=1e-13)
    p.close()


def test_constant_loglik_is_one(dev):
    import torch
    import vsmc_amd as va
    n = 70001
    p = va.Particle(n, 1, "float64", seed=3)
    p.set_log(torch.randn(n, dtype=torch.float64, device="cuda"))
    ll = torch.full((n,), -2.5, dtype=torch.float64, device="cuda")
    for d in (0.1, 1.0, 10.0):
        assert p.cess(ll, d) == pytest.approx(1.0, rel=1e-13)
    assert p.cess_solve(ll, 0.9, 7.0) == 7.0
    p.close()


@pytest.mark.parametrize("target", [0.5, 0.9, 0.99])
@pytest.mark.parametrize("dmax", [0.2, 1.0, 10.0])
def test_solve_vs_oracle(dev, target, dmax):
    import torch
    import vsmc_amd as va
    n = 200003
    p, llt, ll, W = _setup(va, torch, n, 5, False, 3)
    rounds = 12
    got = p.cess_solve(llt, target, dmax, rounds)
    ref = ol.cess_solve(W, ll, dmax, target, tol=1e-15 * dmax)
    print(f"target={target} dmax={dmax} d={got!r} oracle={ref!r}", flush=True)
    if ol.cess(W, ll, dmax) >= target:
        assert got == dmax
    else:
        assert 0.0 <= got < dmax
        assert abs(got - ref) <= 2 * dmax / 16.0 ** rounds + 1e-13 * dmax
        # the returned increment keeps the CESS at the target
        assert ol.cess(W, ll, got) >= target * (1 - 1e-12)
    p.close()


def test_tempering_loop(dev):
    """A tempered sequence from the prior to the posterior of a conjugate
    normal model: each step takes the largest increment that keeps CESS at
    0.9, reweights by d ll, resamples when the ESS halves; the final
    weighted mean agrees with the closed-form posterior mean."""
    import torch
    import vsmc_amd as va
    n, seed = 1 << 18, 2024
    rs = np.random.default_rng(seed)
    data = rs.normal(1.3, 1.0, 20)
    # prior N(0, 3^2) on theta, particles drawn from it
    theta = torch.as_tensor(rs.normal(0, 3.0, n), device="cuda")
    p = va.Particle(n, 1, "float64", seed=seed)
    p.set_state(theta.reshape(1, n))
    p.set_equal()
    alpha, steps = 0.0, 0
    while alpha < 1.0 and steps < 200:
        x = p.state().reshape(n)
        ll = -0.5 * ((torch.as_tensor(data, device="cuda")[None, :] - x[:, None]) ** 2).sum(1)
        d = p.cess_solve(ll, 0.9, 1.0 - alpha)
        assert d > 0
        p.add_log(d * ll)
        alpha = alpha + d if d < 1.0 - alpha else 1.0
        p.resample(ol.SYSTEMATIC, 0.5)
        steps += 1
    assert alpha == 1.0
    W = p.weights()
    x = p.state().reshape(n)
    post_var = 1.0 / (1.0 / 9.0 + len(data))
    post_mean = post_var * data.sum()
    est = float((W * x).sum())
    print(f"steps={steps} posterior mean {post_mean:.5f} estimate {est:.5f}", flush=True)
    assert abs(est - post_mean) < 5 * np.sqrt(post_var / n * 20)
    p.close()


def test_invalid_arguments(dev):
    import torch
    import vsmc_amd as va
    p = va.Particle(100, 1, "float64", seed=1)
    ll = torch.zeros(100, dtype=torch.float64, device="cuda")
    for args in ((0.0, 1.0, 4), (1.5, 1.0, 4), (0.5, -1.0, 4), (0.5, float("inf"), 4),
                 (0.5, 1.0, 0), (0.5, 1.0, 65)):
        with pytest.raises(va.VsmchError):
            p.cess_solve(ll, *args)
    with pytest.raises(va.VsmchError):
        p.cess(ll, -1.0)
    p.close()
