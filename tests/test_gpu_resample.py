"""HIP resampling against the oracle: exact cumulative sums and replication
counts bit-exact (the sequential FP64 fold of transform.hpp:61-69),
ancestors bit-exact, stream word consumption exact; sorted uniforms (libm
log/exp) within 1e-12."""
import math

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def _np(t):
    return t.cpu().numpy()


def weights(kind, n, seed):
    rs = np.random.default_rng(seed)
    if kind == "uniform":
        w = rs.random(n)
    elif kind.startswith("lognormal"):
        s = float(kind.split(":")[1])
        lw = rs.normal(0, s, n)
        w = np.exp(lw - lw.max())
    elif kind == "equal":
        w = np.ones(n)
    elif kind == "onehot":
        w = np.zeros(n)
        w[rs.integers(0, n)] = 1.0
    elif kind == "zeros_head":
        w = rs.random(n)
        w[: n // 3] = 0.0
    elif kind == "tiny_head":
        w = rs.random(n)
        w[:50] = np.ldexp(rs.random(min(50, n)), -1070)
    elif kind == "ties":
        w = rs.choice([1.0, 2.0 ** -53, 2.0 ** -52, 3 * 2.0 ** -53, 2.0 ** -54], n)
    else:
        raise ValueError(kind)
    if kind not in ("ties", "tiny_head") and w.sum() > 0:
        w = w / w.sum()
    return w


KINDS = ["uniform", "lognormal:0.5", "lognormal:2", "lognormal:8", "equal", "onehot",
         "zeros_head", "tiny_head", "ties"]
SIZES = [1, 2, 3, 4095, 4096, 4097, 65539, 1 << 18]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", SIZES)
def test_exact_cumsum(dev, kind, n):
    import vsmc_amd as va
    print(f"kind={kind} n={n}", flush=True)
    w = weights(kind, n, n)
    got = _np(va.exact_cumsum(_t(w)))
    exp = ol.cumsum(w)
    bad = np.nonzero(got != exp)[0]
    assert len(bad) == 0, (bad[:5], got[bad[:3]], exp[bad[:3]])


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [2, 3, 1000, 4097, 65539, 1 << 18])
def test_trans_u01_rep(dev, kind, n):
    import vsmc_amd as va
    print(f"kind={kind} n={n}", flush=True)
    W = weights(kind, n, 7 * n)
    rs = np.random.default_rng(n)
    for N in (n, n // 2 + 1, 2 * n):
        u = np.sort(rs.random(N))
        got = _np(va.resample_trans_u01_rep(_t(W), _t(u)))
        exp = ol.trans_u01_rep(W, u)
        assert (got == exp).all(), (N, np.nonzero(got != exp)[0][:5])
    # unsorted uniforms follow the merge loop too
    u = rs.random(n)
    assert (_np(va.resample_trans_u01_rep(_t(W), _t(u))) == ol.trans_u01_rep(W, u)).all()


def test_trans_u01_rep_boundaries(dev):
    """u placed exactly at, and one ulp around, sequential cumsums."""
    import vsmc_amd as va
    rs = np.random.default_rng(3)
    n = 20000
    W = weights("lognormal:2", n, 3)
    s = ol.cumsum(W)
    pick = rs.integers(0, n - 1, 3000)
    u = np.sort(np.concatenate([s[pick], np.nextafter(s[pick], 0),
                                np.nextafter(s[pick], 2), rs.random(3000)]))
    got = _np(va.resample_trans_u01_rep(_t(W), _t(u)))
    assert (got == ol.trans_u01_rep(W, u)).all()


def test_trans_edges(dev):
    import vsmc_amd as va
    W = _t(np.array([0.25, 0.25, 0.5]))
    assert _np(va.resample_trans_u01_rep(W, _t(np.array([0.25, 0.5, 0.5, 0.75])))).tolist() == [1, 2, 1]
    assert _np(va.resample_trans_u01_rep(_t(np.array([1.0])), _t(np.array([0.1, 0.9])))).tolist() == [2]
    assert _np(va.resample_trans_u01_rep(W, _t(np.zeros(1)), n=0)).tolist() == [0, 0, 0]
    with pytest.raises(va.VsmchError):
        va.resample_trans_u01_rep(_t(np.array([0.5, -0.1, 0.6])), _t(np.array([0.3])))
    with pytest.raises(va.VsmchError):
        va.resample_trans_u01_rep(_t(np.array([0.5, math.nan, 0.5])), _t(np.array([0.3])))


@pytest.mark.parametrize("n", [1, 2, 17, 4096, 4097, 100003])
@pytest.mark.parametrize("kind", ["uniform", "lognormal:2", "lognormal:8", "onehot"])
def test_rep_index(dev, n, kind):
    import torch
    import vsmc_amd as va
    print(f"kind={kind} n={n}", flush=True)
    W = weights(kind, n, n + 1)
    rs = np.random.default_rng(n)
    rep = ol.trans_u01_rep(W, np.sort(rs.random(n)))
    idx = _np(va.resample_trans_rep_index(_t(rep, torch.int32)))
    assert (idx == ol.trans_rep_index(rep)).all()
    back = _np(va.resample_trans_index_rep(_t(idx, torch.int32), n))
    assert (back == rep).all()
    # M != N: expanded list
    N2 = n + 5
    rep2 = ol.trans_u01_rep(W, np.sort(rs.random(N2)))
    idx2 = _np(va.resample_trans_rep_index(_t(rep2, torch.int32), N2))
    assert (idx2 == ol.trans_rep_index(rep2, N2)).all()
    u = np.sort(rs.random(N2))
    assert (_np(va.resample_trans_u01_index(_t(W), _t(u))) == ol.trans_u01_index(W, u)).all()


@pytest.mark.parametrize("n", [1, 5, 4097, 100003])
def test_residual(dev, n):
    import vsmc_amd as va
    W = weights("lognormal:2", n, 5)
    N = n + 3
    resid, integ, R = va.resample_trans_residual(_t(W), N)
    eresid, einteg, eR = ol.trans_residual(W, N)
    assert R == eR
    assert (_np(integ) == einteg).all()
    # M = 1: every residual is 0, 1 / 0 * 0 is NaN in both (as in the reference)
    np.testing.assert_array_equal(_np(resid), eresid)


@pytest.mark.parametrize("scheme", range(6))
@pytest.mark.parametrize("n", [2, 10000, 1 << 20])
def test_schemes_stream(dev, scheme, n):
    """Counts and consumed words of Resample<Scheme> with the Philox
    resampling engine, two events in a row (the second continues the
    stream)."""
    import vsmc_amd as va
    print(f"scheme={scheme} n={n}", flush=True)
    key = 4242
    off = 0
    for ev in range(2):
        W = weights("lognormal:2" if ev == 0 else "uniform", n, n + ev)
        rep, off2 = va.resample(scheme, _t(W), n, key, off)
        rep = _np(rep)
        erep, used = ol.resample(scheme, W, n, key, skip=off)
        if scheme in (ol.MULTINOMIAL, ol.RESIDUAL):
            # libm log/exp inside the sorted uniforms: exact given the GPU's
            # uniforms, checked below; here equal counts are the norm
            mism = int((rep != erep).sum())
            assert mism <= 2, mism
        else:
            assert (rep == erep).all(), np.nonzero(rep != erep)[0][:5]
        assert rep.sum() == n
        if scheme not in (ol.MULTINOMIAL, ol.RESIDUAL):
            assert off2 - off == used
        off = off2


@pytest.mark.parametrize("n", [1, 2, 1000, 65539])
def test_sorted_uniforms(dev, n):
    """U01SequenceSorted on the device vs the oracle (libm log/exp: 1e-12),
    and the counts from those device uniforms are bit-exact."""
    import vsmc_amd as va
    key, off = 77, 5
    got = _np(va.sorted_uniforms(n, key, off))
    exp = ol.scheme_u01(ol.MULTINOMIAL, n, key, n, skip=off)
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-12)
    W = weights("lognormal:2", n, 1)
    rep, off2 = va.resample(ol.MULTINOMIAL, _t(W), n, key, off)
    vmax = np.maximum.accumulate(got)
    assert (_np(rep) == ol.trans_u01_rep(W, vmax)).all()


@pytest.mark.parametrize("scheme", [ol.SYSTEMATIC, ol.STRATIFIED])
@pytest.mark.parametrize("s", [0.5, 2.0, 8.0])
def test_full_size_counts(dev, scheme, s):
    """N = 2^24 log-weights ~ N(0, s^2): counts bit-exact with the oracle's
    sequential merge; sum = N."""
    import vsmc_amd as va
    n = 1 << 24
    print(f"scheme={scheme} s={s} n={n}", flush=True)
    rs = np.random.default_rng(int(s * 10))
    lw = rs.normal(0, s, n)
    W = np.exp(lw - lw.max())
    W /= W.sum()
    rep, _ = va.resample(scheme, _t(W), n, 99, 0)
    rep = _np(rep)
    erep, _ = ol.resample(scheme, W, n, 99)
    assert rep.sum() == n
    bad = np.nonzero(rep != erep)[0]
    assert len(bad) == 0, (len(bad), bad[:5])


@pytest.mark.parametrize("s", [0.5, 2.0, 8.0])
def test_full_size_multinomial_residual(dev, s):
    """N = 2^24 log-weights ~ N(0, s^2), the schemes whose uniforms need libm
    log / exp (multinomial: U01SequenceSorted, u01_sequence.hpp:130-164;
    residual: the residuals resampled with it, residual.hpp:43-58). Counts
    bit-exact with the oracle's merge given the device's own uniforms; against
    the oracle's libm uniforms a handful of counts may move. The residual
    stratified / systematic schemes are bit-exact with the oracle."""
    import vsmc_amd as va
    n = 1 << 24
    rs = np.random.default_rng(int(s * 100))
    lw = rs.normal(0, s, n)
    W = np.exp(lw - lw.max())
    W /= W.sum()
    Wd = _t(W)
    key, off = 123, 7
    print(f"multinomial s={s} n={n}", flush=True)
    rep, off2 = va.resample(ol.MULTINOMIAL, Wd, n, key, off)
    rep = _np(rep)
    u = _np(va.sorted_uniforms(n, key, off))
    assert (np.diff(u) >= 0).all()
    assert (rep == ol.trans_u01_rep(W, u)).all()
    erep, _ = ol.resample(ol.MULTINOMIAL, W, n, key, skip=off)
    mism = int((rep != erep).sum())
    print("multinomial counts moved by libm:", mism, flush=True)
    assert mism <= 64
    assert rep.sum() == n
    print(f"residual s={s}", flush=True)
    rep, _ = va.resample(ol.RESIDUAL, Wd, n, key, off)
    rep = _np(rep)
    resid, integ, R = va.resample_trans_residual(Wd, n)
    eresid, einteg, eR = ol.trans_residual(W, n)
    assert R == eR and (_np(integ) == einteg).all()
    assert (_np(resid) == eresid).all()
    ur = _np(va.sorted_uniforms(R, key, off))
    assert (rep == einteg + ol.trans_u01_rep(eresid, ur)).all()
    assert rep.sum() == n
    for scheme in (ol.RESIDUAL_STRATIFIED, ol.RESIDUAL_SYSTEMATIC):
        print(f"scheme={scheme} s={s}", flush=True)
        rep, off3 = va.resample(scheme, Wd, n, key, off)
        erep, used = ol.resample(scheme, W, n, key, skip=off)
        rep = _np(rep)
        bad = np.nonzero(rep != erep)[0]
        assert len(bad) == 0, (scheme, len(bad), bad[:5])
        assert off3 - off == used


def test_growing_weights_counts(dev):
    """Weights growing by a factor 2 every 2^18 particles (N = 2^24): the
    running sum crosses a binade in every group of 64 tiles, more crossing
    groups and tiles than the resolver holds in LDS, so the rest are staged
    as the walk reaches them. Counts bit-exact with the oracle."""
    import vsmc_amd as va
    n = 1 << 24
    i = np.arange(n, dtype=np.float64)
    rs = np.random.default_rng(3)
    W = np.exp2(i / (1 << 18)) * (0.5 + rs.random(n))
    W /= W.sum()
    for scheme in (ol.SYSTEMATIC, ol.STRATIFIED):
        print(f"scheme={scheme} n={n} growing", flush=True)
        rep, _ = va.resample(scheme, _t(W), n, 17, 0)
        rep = _np(rep)
        erep, _ = ol.resample(scheme, W, n, 17)
        bad = np.nonzero(rep != erep)[0]
        assert len(bad) == 0, (len(bad), bad[:5])


def test_several_resolver_windows(dev):
    """N = 2^26 + 12345 (16,388 tiles, 257 groups): the resolver stages its
    summaries in five windows of 64 groups. Counts bit-exact with the oracle."""
    import vsmc_amd as va
    n = (1 << 26) + 12345
    print(f"n={n} systematic", flush=True)
    rs = np.random.default_rng(26)
    W = np.exp(rs.normal(0, 1.0, n))
    W /= W.sum()
    rep, _ = va.resample(ol.SYSTEMATIC, _t(W), n, 5, 0)
    rep = _np(rep)
    erep, _ = ol.resample(ol.SYSTEMATIC, W, n, 5)
    assert rep.sum() == n
    bad = np.nonzero(rep != erep)[0]
    assert len(bad) == 0, (len(bad), bad[:5])
