"""Device ResampleIndex (vsmch_rindex_*) against the oracle: random
ancestor rows and identity rows, insert at the last and at an earlier
iteration, reset / clear, both matrix layouts, index(id, iter), a history
fed by a particle filter's own resampling events, and the error (no
device access) for an ancestor outside [0, N)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, np.int32), device="cuda")


def test_history_vs_oracle(dev):
    import vsmc_amd as va
    N, T = 3000, 25
    rs = np.random.default_rng(1)
    r = va.ResampleIndex(N)
    hist = []
    for t in range(T):
        if t % 3 == 1:
            r.push_back()
            hist.append(np.arange(N))
        else:
            row = rs.integers(0, N, N)
            r.push_back(_t(row))
            hist.append(row)
    assert r.iter_size() == T
    h = np.array(hist, np.int64)
    print(f"N={N} T={T}", flush=True)
    np.testing.assert_array_equal(r.index_matrix(va.ROW_MAJOR).cpu().numpy(),
                                  ol.rindex_matrix(h, 101))
    np.testing.assert_array_equal(r.index_matrix(va.COL_MAJOR).cpu().numpy(),
                                  ol.rindex_matrix(h, 102))
    for i, t in [(0, 0), (17, 3), (N - 1, T - 1), (1234, 10)]:
        assert r.index(i, t) == ol.rindex_index(h, i, t)
    # insert: overwrite the last row and an earlier one
    row = rs.integers(0, N, N)
    r.insert(_t(row))
    h[T - 1] = row
    r.insert(None, 4)
    h[4] = np.arange(N)
    np.testing.assert_array_equal(r.index_matrix().cpu().numpy(), ol.rindex_matrix(h, 101))
    # growth across capacity, then reset and clear
    for _ in range(40):
        r.push_back()
    assert r.iter_size() == T + 40
    r.reset()
    assert r.iter_size() == 0
    r.push_back(_t(h[0]))
    np.testing.assert_array_equal(r.index_matrix().cpu().numpy()[:, 0], h[0])
    r.clear()
    assert r.iter_size() == 0
    r.close()


def test_pf_lineages(dev):
    """Ancestors of a constant-velocity filter's resampling events: the
    device history reproduces the composition of the events' idx arrays."""
    import vsmc_amd as va
    from vsmc_amd.models import cv_simulate
    n, T = 1 << 14, 30
    _, ox, oy = cv_simulate(T, 7)
    f = va.CvFilter(n, "float32", seed=4)
    f.init(float(ox[0]), float(oy[0]))
    r = va.ResampleIndex(n)
    r.push_back()
    rows = [np.arange(n)]
    nres = 0
    for t in range(1, T):
        f.move(t, float(ox[t]), float(oy[t]))
        did, _, idx = f.resample(va.SYSTEMATIC, 0.5, want_rep=True)
        if did:
            nres += 1
            r.push_back(idx)
            rows.append(idx.cpu().numpy().astype(np.int64))
        else:
            r.push_back()
            rows.append(np.arange(n))
    f.close()
    print(f"n={n} T={T} resampled={nres}", flush=True)
    assert nres >= 3
    h = np.array(rows)
    m = r.index_matrix(va.ROW_MAJOR).cpu().numpy()
    np.testing.assert_array_equal(m, ol.rindex_matrix(h, 101))
    # lineages coalesce going back: fewer distinct ancestors at t = 0
    assert len(np.unique(m[:, 0])) < len(np.unique(m[:, -1]))
    r.close()


def test_out_of_range_ancestor_is_an_error(dev):
    """An ancestor index outside [0, N) is reported, never dereferenced."""
    import vsmc_amd as va
    N = 100
    r = va.ResampleIndex(N)
    row = np.arange(N)
    row[7] = N + 5
    r.push_back(_t(row))
    r.push_back(_t(np.full(N, 7)))
    with pytest.raises(va.VsmchError):
        r.index_matrix()
    with pytest.raises(va.VsmchError):
        r.index(3, 0)
    assert r.index(3, 1) == 7
    r.close()


def test_cpp_layer_vs_oracle(dev, tmp_path):
    """include/vsmch/resample_index.hpp (tests/cpp/rindex_hip.cpp): both
    matrix layouts and index(i, t) against the oracle."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "rindex_hip")
    N, T = 37, 9
    rs = np.random.default_rng(12)
    rows, lines = [], []
    for t in range(T):
        if t in (2, 5):
            rows.append(np.arange(N))
            lines.append("I")
        else:
            row = rs.integers(0, N, N)
            rows.append(row)
            lines.append(" ".join(str(int(v)) for v in row))
    path = tmp_path / "hist.txt"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([exe, str(N), str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = [[int(x) for x in ln.rstrip("\t").split("\t")] for ln in r.stdout.strip().split("\n")]
    h = np.array(rows, np.int64)
    ref = ol.rindex_matrix(h, 101)
    np.testing.assert_array_equal(np.array(out[:N]), ref)
    np.testing.assert_array_equal(np.array(out[N:N + T]), ref.T)
    np.testing.assert_array_equal(np.array(out[N + T:]), ref)
