"""The oracle's Sampler loop (vo_pf_run) against an independent numpy
restatement of the reference's step order, for T = 3:

  initialize: init functor, weights set_log, then do_resample
  iterate:    move functor, weights add_log, then do_resample
  do_resample records ESS *before* Particle::resample, resamples when
  ESS < threshold * N, and the monitor runs after it
  (include/vsmc/core/sampler.hpp:735-785, particle.hpp:170-197,
  monitor.hpp:242-268).

The restatement builds the normals from the raw Philox words (Box-Muller of
normal_distribution.hpp:199-216), the systematic counts from a sequential
numpy cumsum (transform.hpp:42-71, u01_sequence.hpp:225-259) and the
rep -> index layout of transform.hpp:116-133, with no call into the oracle
except the Philox words. Wrong orders planted in the restatement (ESS after
resampling, monitor before it, the decision from the previous step's ESS)
must each disagree with vo_pf_run, so the comparison holds the order."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SEED = 20161


def c0_obs():
    with open(os.path.join(GOLD, "cv_obs_c0.json")) as f:
        g = json.load(f)
    return (np.array([float.fromhex(x) for x in g["obs_x"]]),
            np.array([float.fromhex(x) for x in g["obs_y"]]))


def _u01(words):
    return words.astype(np.float64) * 2.0 ** -32


def _normals(key, block):
    """The four normals of one init/move of particle `key`: two scalar
    Box-Muller pairs (one NormalDistribution per pair of components, a
    fresh object per eval_sp), words 4 block .. 4 block + 3."""
    u = _u01(ol.philox_words(key, 4 * block, 4))
    r0, a0 = math.sqrt(-2 * math.log(u[0])), 2 * math.pi * u[1]
    r1, a1 = math.sqrt(-2 * math.log(u[2])), 2 * math.pi * u[3]
    return r0 * math.cos(a0), r0 * math.sin(a0), r1 * math.cos(a1), r1 * math.sin(a1)


def _loglik(x, y, ox, oy):
    rx, ry = (x - ox) / 0.1, (y - oy) / 0.1
    return -5.5 * (np.log(1 + rx * rx / 10) + np.log(1 + ry * ry / 10))


def _systematic_counts(W, U):
    """Counts of the sequential merge with u_j = U delta + j delta."""
    N = len(W)
    delta = 1.0 / N
    u0 = U * delta
    u = np.array([u0 + j * delta for j in range(N)])
    acc = np.cumsum(W)                       # left-to-right, as accumulate
    le = np.searchsorted(u, acc, side="right")
    rep = np.diff(np.concatenate([[0], le]))
    rep[-1] = N - le[-2] if N > 1 else N
    return rep


def _rep_index(rep):
    N = len(rep)
    idx = np.arange(N)
    zeros = np.nonzero(rep == 0)[0]
    extra = np.repeat(np.arange(N), np.maximum(rep - 1, 0))
    idx[zeros] = extra
    return idx


def restated_run(N, T, threshold, ox, oy, variant="reference"):
    keys = SEED + 1 + np.arange(N)
    rkey = SEED + N + 1
    rword = 0
    x = np.zeros((N, 4))
    W = np.full(N, 1.0 / N)
    ess_prev = N
    out = {"ess": np.zeros(T), "resampled": np.zeros(T, np.int64),
           "pos": np.zeros((T, 2)), "logz": np.zeros(T)}
    logz = 0.0
    for t in range(T):
        for i in range(N):
            z = _normals(int(keys[i]), t)
            if t == 0:
                x[i] = [2.0 * z[0], 2.0 * z[1], 1.0 * z[2], 1.0 * z[3]]
            else:
                x[i, 0] += z[0] * 0.02 + 0.1 * x[i, 2]
                x[i, 1] += z[1] * 0.02 + 0.1 * x[i, 3]
                x[i, 2] += z[2] * 0.02
                x[i, 3] += z[3] * 0.02
        ll = _loglik(x[:, 0], x[:, 1], ox[t], oy[t])
        lw = (np.log(W) if t > 0 else 0.0) + ll
        m = lw.max()
        logz += m + math.log(np.exp(lw - m).sum()) + (math.log(1.0 / N) if t == 0 else 0.0)
        W = np.exp(lw - m)
        W /= W.sum()
        ess = 1.0 / (W @ W)
        decide = ess_prev if variant == "stale_decision" else ess
        if variant == "monitor_first":
            out["pos"][t] = W @ x[:, :2]
        do = decide < threshold * N
        if do:
            U = _u01(ol.philox_words(rkey, rword, 1))[0]
            rword += 1
            idx = _rep_index(_systematic_counts(W, U))
            x = x[idx]
            W = np.full(N, 1.0 / N)
        ess_after = 1.0 / (W @ W)
        out["ess"][t] = ess_after if variant == "ess_after" else ess
        out["resampled"][t] = int(do)
        if variant != "monitor_first":
            out["pos"][t] = W @ x[:, :2]
        out["logz"][t] = logz
        ess_prev = ess
    return out, x, W


def _agree(a, b):
    return ((a["resampled"] == b["resampled"]).all()
            and np.allclose(a["ess"], b["ess"], rtol=1e-10, atol=0)
            and np.allclose(a["pos"], b["pos"], rtol=1e-10, atol=1e-12)
            and np.allclose(a["logz"], b["logz"], rtol=1e-10, atol=1e-12))


CASES = [("always", math.inf), ("never", -math.inf), ("half", 0.5)]


@pytest.mark.parametrize("name,thr", CASES)
def test_pf_order_matches_restatement(name, thr):
    ox, oy = c0_obs()
    N, T = 48, 3
    got = ol.pf_run(N, T, SEED, ol.SYSTEMATIC, thr, ox, oy, want_state=True)
    exp, x, W = restated_run(N, T, thr, ox, oy)
    print(name, got["resampled"], got["ess"])
    assert (got["resampled"] == exp["resampled"]).all(), (got["resampled"], exp["resampled"])
    np.testing.assert_allclose(got["ess"], exp["ess"], rtol=1e-10)
    np.testing.assert_allclose(got["pos"], exp["pos"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(got["logz"], exp["logz"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(got["state"], x, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got["W"], W, rtol=1e-12)
    if name == "always":
        # ESS is the pre-resampling value, not the N of the equal weights
        assert (got["ess"] < 0.9 * N).all()
        # after a resampling step the monitor is the plain mean of the state
        np.testing.assert_allclose(got["pos"][-1], got["state"][:, :2].mean(0), rtol=1e-12)
    if name == "half":
        # the mixed case: some steps resample, some do not
        assert 0 < got["resampled"].sum() < T


@pytest.mark.parametrize("variant", ["ess_after", "monitor_first", "stale_decision"])
def test_planted_wrong_orders_disagree(variant):
    """Each wrong order fails against vo_pf_run in at least one threshold
    case, so the comparison above would catch it in the oracle."""
    ox, oy = c0_obs()
    N, T = 48, 3
    caught = False
    for _, thr in CASES:
        got = ol.pf_run(N, T, SEED, ol.SYSTEMATIC, thr, ox, oy)
        bad, _, _ = restated_run(N, T, thr, ox, oy, variant=variant)
        caught = caught or not _agree(got, bad)
    assert caught, variant
