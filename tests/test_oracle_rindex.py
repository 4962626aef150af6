"""The oracle's ResampleIndex (index(id, iter) and index_matrix, both
layouts) against direct composition of the ancestor rows: the lineage of
particle i at iteration t is h_t[h_{t+1}[... h_{T-1}[i]]]."""
import numpy as np
import pytest

import oracle_lib as ol


def _compose(hist):
    T, N = hist.shape
    out = np.empty((N, T), np.int64)
    cur = np.arange(N)
    for t in range(T - 1, -1, -1):
        cur = hist[t][cur]
        out[:, t] = cur
    return out


def _hist(rs, N, T, p_identity=0.4):
    rows = []
    for _ in range(T):
        if rs.random() < p_identity:
            rows.append(np.arange(N))
        else:
            rows.append(rs.integers(0, N, N))
    return np.array(rows, np.int64)


@pytest.mark.parametrize("N,T", [(1, 1), (5, 1), (7, 3), (50, 17), (1000, 40)])
def test_matrix_both_layouts(N, T):
    rs = np.random.default_rng(N * 100 + T)
    h = _hist(rs, N, T)
    ref = _compose(h)
    np.testing.assert_array_equal(ol.rindex_matrix(h, 101), ref)
    np.testing.assert_array_equal(ol.rindex_matrix(h, 102), ref.T)


def test_index_matches_matrix():
    rs = np.random.default_rng(3)
    h = _hist(rs, 64, 12)
    m = ol.rindex_matrix(h, 101)
    for i in range(0, 64, 7):
        for t in range(12):
            assert ol.rindex_index(h, i, t) == m[i, t]


def test_identity_history():
    """Steps without resampling (identity rows) leave every lineage fixed;
    the last row alone decides the last column."""
    N, T = 33, 6
    h = np.tile(np.arange(N), (T, 1))
    np.testing.assert_array_equal(ol.rindex_matrix(h, 101), np.tile(np.arange(N)[:, None], (1, T)))
    h[T - 1] = np.roll(np.arange(N), 5)
    m = ol.rindex_matrix(h, 101)
    np.testing.assert_array_equal(m[:, T - 1], h[T - 1])
    np.testing.assert_array_equal(m[:, 0], h[T - 1])


def test_plant_wrong_order_fails():
    """Composing the rows first-to-last instead of last-to-first (a plausible
    mistake) differs from the oracle."""
    rs = np.random.default_rng(9)
    h = _hist(rs, 40, 5, p_identity=0.0)
    wrong = np.empty((40, 5), np.int64)
    cur = np.arange(40)
    for t in range(5):
        cur = h[t][cur]
        wrong[:, t] = cur
    assert not np.array_equal(ol.rindex_matrix(h, 101), wrong)
