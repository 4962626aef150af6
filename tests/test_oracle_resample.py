"""The oracle's resampling restatement pinned by brute force, invariants and
closed forms (the reference's C++ does not build here: it needs CBLAS and
LAPACKE headers the image lacks).

Brute force: the count of particle i is the number of u_j (in order, up to
the first one above the running cumulative weight) that fall at or below the
Python left-fold cumulative sum; computed here with an explicit "first j with
u_j > s" search that shares no code with the oracle's merge loop.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol


def brute_counts(W, u):
    """rep_i = G(s_i) - G(s_{i-1}), G(x) = first j with max(u_0..u_j) > x."""
    M, N = len(W), len(u)
    if M == 0:
        return np.zeros(0, np.int64)
    if M == 1:
        return np.array([N], np.int64)
    s = []
    acc = 0.0
    for w in W[:-1]:
        acc = acc + float(w)
        s.append(acc)

    def G(x):
        m = -math.inf
        for j, v in enumerate(u):
            m = max(m, v)
            if m > x:
                return j
        return N

    g = [G(x) for x in s]
    rep = [g[0]] + [g[i] - g[i - 1] for i in range(1, M - 1)] + [N - g[-1]]
    return np.array(rep, np.int64)


def brute_rep_index(rep):
    """k-th childless slot (dst order) takes the source whose extra copies
    (rep - 1, src order) cover k (transform.hpp:116-133)."""
    N = len(rep)
    extras = [s for s in range(N) for _ in range(max(rep[s] - 1, 0))]
    idx = np.zeros(N, np.int64)
    k = 0
    for d in range(N):
        if rep[d] > 0:
            idx[d] = d
        else:
            idx[d] = extras[k]
            k += 1
    return idx


@pytest.mark.parametrize("seed", range(40))
def test_trans_u01_rep_brute(seed):
    rs = np.random.default_rng(seed)
    M = int(rs.integers(0, 40))
    N = int(rs.integers(0, 40))
    kind = seed % 4
    if kind == 0:
        W = rs.random(M)
    elif kind == 1:
        W = np.exp(rs.normal(0, 3, M))
    elif kind == 2:
        W = rs.choice([0.0, 0.25, 0.5, 1.0 / 3], size=M)
    else:
        W = np.zeros(M)
        if M:
            W[rs.integers(0, M)] = 1.0
    if M and W.sum() > 0:
        W = W / W.sum()
    u = np.sort(rs.random(N)) if seed % 5 else rs.random(N)   # unsorted too
    got = ol.trans_u01_rep(W, u)
    exp = brute_counts(W, u)
    assert (got == exp).all()
    if M:
        assert got.sum() == N


def test_trans_u01_rep_ties_and_edges():
    # u exactly equal to a cumulative sum counts for that particle (<=)
    W = np.array([0.25, 0.25, 0.5])
    u = np.array([0.25, 0.5, 0.5, 0.75])
    assert ol.trans_u01_rep(W, u).tolist() == [1, 2, 1]
    assert ol.trans_u01_rep(np.array([1.0]), u).tolist() == [4]
    assert ol.trans_u01_rep(W, np.zeros(0)).tolist() == [0, 0, 0]
    assert ol.trans_u01_rep(np.zeros(0), u).tolist() == []
    # the last particle takes everything beyond s_{M-2}
    assert ol.trans_u01_rep(np.array([0.0, 0.0]), u).tolist() == [0, 4]


def test_sequential_fold_matters():
    """Counts follow the left fold, not the exact sum: 0.1 + 0.2 rounds up."""
    W = np.array([0.1, 0.2, 0.7])
    s1 = 0.1 + 0.2
    u = np.array([np.nextafter(s1, 0), s1])
    assert ol.trans_u01_rep(W, u).tolist() == [0, 2, 0]


@pytest.mark.parametrize("seed", range(30))
def test_rep_index_brute(seed):
    rs = np.random.default_rng(100 + seed)
    N = int(rs.integers(1, 60))
    W = np.exp(rs.normal(0, 1 + seed % 3, N))
    W /= W.sum()
    rep = ol.trans_u01_rep(W, np.sort(rs.random(N)))
    idx = ol.trans_rep_index(rep)
    assert (idx == brute_rep_index(rep)).all()
    assert ((rep > 0) <= (idx == np.arange(N))).all()
    # index_rep inverts it
    back = np.zeros(N, np.int64)
    ol.oracle().vo_trans_index_rep(N, N, idx, back)
    assert (back == rep).all()
    # M != N: expanded list in source order
    M = N
    Nx = int(rs.integers(1, 80))
    if Nx == M:
        Nx += 1
    rep2 = ol.trans_u01_rep(W, np.sort(rs.random(Nx)))
    idx2 = ol.trans_rep_index(rep2, Nx)
    assert (idx2 == np.repeat(np.arange(M), rep2)).all()
    assert (ol.trans_u01_index(W, np.sort(rs.random(Nx))) >= 0).all()


def test_u01_index_matches_expansion():
    rs = np.random.default_rng(5)
    W = rs.random(30)
    W /= W.sum()
    u = np.sort(rs.random(50))
    rep = ol.trans_u01_rep(W, u)
    assert (ol.trans_u01_index(W, u) == np.repeat(np.arange(30), rep)).all()


def test_residual_closed_form():
    rs = np.random.default_rng(9)
    W = rs.random(50)
    W /= W.sum()
    N = 73
    resid, integ, R = ol.trans_residual(W, N)
    frac = np.array([math.modf(N * w)[0] for w in W])
    ip = np.array([int(math.modf(N * w)[1]) for w in W])
    acc = 0.0
    for f in frac:
        acc += f
    assert (integ == ip).all()
    assert R == N - ip.sum()
    assert (resid == (1 / acc) * frac).all()


def test_scheme_sequences_closed_form():
    key = 1234
    N = 1000
    U = ol.u01_fill(0, key, 0, N + 1)
    delta = 1 / N
    # stratified: u_j = U_j * delta + j * delta (u01_sequence.hpp:200)
    strat = ol.scheme_u01(ol.STRATIFIED, N, key, N)
    assert (strat == U[:N] * delta + np.arange(N) * delta).all()
    # systematic: u0 = U_0 * delta, u_j = u0 + j * delta (:231-246)
    sysu = ol.scheme_u01(ol.SYSTEMATIC, N, key, N)
    assert (sysu == U[0] * delta + np.arange(N) * delta).all()
    # sorted: lmax += log(U_j) / (N - j); u_j = 1 - exp(lmax) (:149-151)
    srt = ol.scheme_u01(ol.MULTINOMIAL, N, key, N)
    lmax = 0.0
    exp = []
    for j in range(N):
        lmax += math.log(U[j]) / (N - j)
        exp.append(1 - math.exp(lmax))
    np.testing.assert_allclose(srt, exp, rtol=0, atol=1e-15)
    assert (np.diff(srt) >= 0).all()


@pytest.mark.parametrize("scheme", range(6))
def test_schemes_words_and_invariants(scheme):
    rs = np.random.default_rng(scheme)
    N = 500
    W = np.exp(rs.normal(0, 2, N))
    W /= W.sum()
    rep, used = ol.resample(scheme, W, N, 777)
    assert rep.sum() == N and (rep >= 0).all()
    if scheme in (ol.SYSTEMATIC, ol.RESIDUAL_SYSTEMATIC):
        assert used == 1
    if scheme == ol.SYSTEMATIC:
        assert (np.abs(rep - N * W) < 1 + 1e-9).all()
    if scheme >= ol.RESIDUAL:
        assert (rep >= np.floor(N * W) - 1e-9).all()
    if scheme in (ol.MULTINOMIAL, ol.STRATIFIED):
        # words consumed: up to and including the first u above s_{M-2}
        jend = N - rep[-1]
        assert used == min(jend + 1, N)
    # counts from the lazily drawn sequence == counts from the materialised one
    nd = N
    if scheme >= ol.RESIDUAL:
        resid, integ, nd = ol.trans_residual(W, N)
        base = {ol.RESIDUAL: ol.MULTINOMIAL, ol.RESIDUAL_STRATIFIED: ol.STRATIFIED,
                ol.RESIDUAL_SYSTEMATIC: ol.SYSTEMATIC}[scheme]
        u = ol.scheme_u01(base, nd, 777, nd)
        assert (ol.trans_u01_rep(resid, u) + integ == rep).all()
    else:
        u = ol.scheme_u01(scheme, N, 777, N)
        assert (ol.trans_u01_rep(W, u) == rep).all()


def test_scheme_edge_sizes():
    rep, used = ol.resample(ol.STRATIFIED, np.array([1.0]), 5, 3)
    assert rep.tolist() == [5] and used == 0
    rep, used = ol.resample(ol.SYSTEMATIC, np.array([1.0]), 5, 3)
    assert rep.tolist() == [5] and used == 1
    rep, used = ol.resample(ol.STRATIFIED, np.array([0.5, 0.5]), 0, 3)
    assert rep.tolist() == [0, 0] and used == 0


def test_copy_and_monitor_closed_form():
    rs = np.random.default_rng(1)
    N, d = 20, 3
    x = rs.random((N, d))
    W = rs.random(N)
    W /= W.sum()
    rep = ol.trans_u01_rep(W, np.sort(rs.random(N)))
    idx = ol.trans_rep_index(rep)
    xr = x.copy().ravel()
    ol.oracle().vo_copy_rowmajor(N, d, xr, idx)
    assert (xr.reshape(N, d) == x[idx]).all()
    xc = np.ascontiguousarray(x.T).ravel()
    ol.oracle().vo_copy_colmajor(N, d, xc, idx)
    assert (xc.reshape(d, N).T == x[idx]).all()
    r = np.zeros(d)
    ol.oracle().vo_monitor_moments(N, d, np.ascontiguousarray(x).ravel(), W, r)
    np.testing.assert_allclose(r, W @ x, rtol=1e-14)
