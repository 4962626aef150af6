"""vsmch_resample_trans_u01_rep_noalloc (the allocation-free form, scratch in
a file-scope device array) against vsmch_resample_trans_u01_rep: bit-exact
counts on two calls in a row, the second with a smaller input (the scratch
keeps the first call's contents), and a clear error beyond its maximum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _case(m, n, s, seed):
    rs = np.random.default_rng(seed)
    lw = rs.normal(0.0, s, m)
    W = np.exp(lw - lw.max())
    W /= W.sum()
    u = np.sort(rs.random(n))
    return W, u


def test_noalloc_matches_allocating_entry(dev):
    import vsmc_amd as va
    cases = [(va.NOALLOC_MAX, va.NOALLOC_MAX, 2.0, 1),   # largest allowed
             (70001, 50000, 8.0, 2)]                     # then smaller
    for m, n, s, seed in cases:
        print(f"noalloc m={m} n={n} s={s}", flush=True)
        W, u = _case(m, n, s, seed)
        Wd, ud = _t(W), _t(u)
        got = va.resample_trans_u01_rep_noalloc(Wd, ud).cpu().numpy()
        exp = va.resample_trans_u01_rep(Wd, ud).cpu().numpy()
        assert got.sum() == n
        bad = np.nonzero(got != exp)[0]
        assert len(bad) == 0, (m, n, bad[:5])


def test_noalloc_edges(dev):
    """m == 1, n == 0 and a size beyond the maximum (rejected, nothing run)."""
    import torch
    import vsmc_amd as va
    W = _t(np.array([1.0]))
    u = _t(np.array([0.25, 0.5, 0.75]))
    assert va.resample_trans_u01_rep_noalloc(W, u).cpu().tolist() == [3]
    W3 = _t(np.array([0.2, 0.3, 0.5]))
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    assert va.resample_trans_u01_rep_noalloc(W3, e, n=0).cpu().tolist() == [0, 0, 0]
    with pytest.raises(va.VsmchError, match="exceed the fixed scratch"):
        va.resample_trans_u01_rep_noalloc(W3, u, m=va.NOALLOC_MAX + 1, n=3)
