"""The reference's u01 sequence C functions (vsmc_rng_u01_sorted /
_stratified / _systematic, include/vsmc/vsmc.h:76-83) and the per-scheme
resampling entry points (vsmc_resample_<scheme>, :297-308) on the device,
against the oracle. Stratified and systematic are bit-exact (the systematic
sums are the exact sequential ones); sorted is within 1e-13 (device and host
log / exp differ in the last ulp)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _oracle(name, n, arg):
    out = np.zeros(max(n, 1))
    if name == "systematic":
        ol.oracle().vo_u01_systematic(n, float(arg), out)
    else:
        getattr(ol.oracle(), "vo_u01_" + name)(n, np.ascontiguousarray(arg, np.float64), out)
    return out[:n]


@pytest.mark.parametrize("n", [1, 2, 4095, 4097, 100003, (1 << 20) + 7])
def test_stratified(dev, n):
    import torch
    import vsmc_amd as va
    u = np.random.default_rng(n).random(n)
    got = va.u01_stratified(torch.as_tensor(u, device="cuda")).cpu().numpy()
    np.testing.assert_array_equal(got, _oracle("stratified", n, u))


@pytest.mark.parametrize("n", [1, 2, 4095, 4097, 100003, (1 << 22) + 3])
@pytest.mark.parametrize("u", [0.0, 0.3, 0.9999999])
def test_systematic(dev, n, u):
    import vsmc_amd as va
    print(f"systematic n={n} u={u}", flush=True)
    got = va.u01_systematic(n, u).cpu().numpy()
    np.testing.assert_array_equal(got, _oracle("systematic", n, u))


@pytest.mark.parametrize("n", [1, 2, 4095, 4097, 100003])
def test_sorted(dev, n):
    import torch
    import vsmc_amd as va
    u = np.random.default_rng(n + 1).random(n)
    got = va.u01_sorted(torch.as_tensor(u, device="cuda")).cpu().numpy()
    exp = _oracle("sorted", n, u)
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13)
    assert (np.diff(got) >= 0).all()


@pytest.mark.parametrize("n", [5, 9000])
def test_sorted_zero_uniform(dev, n):
    """A uniform of exactly 0 (log 0 = -inf, an infinite term of the fold):
    that value and every later one is 1, as in the reference."""
    import torch
    import vsmc_amd as va
    u = np.random.default_rng(n + 2).random(n)
    u[n // 3] = 0.0
    print(f"sorted with a zero uniform n={n}", flush=True)
    got = va.u01_sorted(torch.as_tensor(u, device="cuda")).cpu().numpy()
    exp = _oracle("sorted", n, u)
    assert (got[n // 3:] == 1.0).all() and (exp[n // 3:] == 1.0).all()
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13)


@pytest.mark.parametrize("idx,name", list(enumerate(
    ["multinomial", "stratified", "systematic", "residual", "residual_stratified",
     "residual_systematic"])))
def test_resample_per_scheme(dev, idx, name):
    """vsmch_resample_<scheme> equals vsmch_resample(..., scheme): counts and
    the engine words consumed, and matches the oracle's Resample<Scheme>."""
    import torch
    import vsmc_amd as va
    n, key, off = 50021, 91, 3
    W = np.random.default_rng(idx).lognormal(0, 1.5, n)
    W /= W.sum()
    Wt = torch.as_tensor(W, device="cuda")
    r1, o1 = va.resample_scheme(name, Wt, n, key, off)
    r2, o2 = va.resample(idx, Wt, n, key, off)
    assert o1 == o2
    np.testing.assert_array_equal(r1.cpu().numpy(), r2.cpu().numpy())
    rep, used = ol.resample(idx, W, n, key, skip=off)
    np.testing.assert_array_equal(r1.cpu().numpy(), rep)
    assert o1 == off + used


def test_invalid(dev):
    import torch
    import vsmc_amd as va
    with pytest.raises(ValueError):
        va.resample_scheme("bogus", torch.ones(4, dtype=torch.float64, device="cuda"), 4, 1)
    with pytest.raises(va.VsmchError):
        va.u01_sorted(torch.as_tensor([0.5, 1.5], dtype=torch.float64, device="cuda"))
