"""Batched filters: every filter of a batch (uneven particle counts and
horizons, its own seed and observations) reproduces, bit for bit, the
single-filter run with the same seed -- log-evidence, ESS and resampling
histories, final state and weights -- with the eager and the captured
HIP-graph step, and with the learned proposal."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def c0_obs():
    with open(os.path.join(GOLD, "cv_obs_c0.json")) as f:
        g = json.load(f)
    return (np.array([float.fromhex(x) for x in g["obs_x"]]),
            np.array([float.fromhex(x) for x in g["obs_y"]]))


# 2^17 and 2^18 - 5 particles: 32 and 64 tiles, so the walk over a filter's
# tile summaries (16 held at a time) carries its cursor and exact start
# across several windows, with the binade crossings of the running sum
# (at 1/4, 1/2 of the total) in later windows
NS = [2, 3, 1000, 4096, 5003, 12288, 40001, 1 << 17, (1 << 18) - 5]
TS = [6, 1, 30, 17, 40, 25, 33, 14, 11]
SEEDS = [7, 8, 9, 10, 11, 12, 13, 14, 15]


def _tracks():
    ox, oy = c0_obs()
    # a different stretch of the track for each filter
    return [(ox[3 * f:], oy[3 * f:]) for f in range(len(NS))]


def _single(n, T, seed, track, scheme, net=None):
    import vsmc_amd as va
    f = va.CvFilter(n, "float32", seed=seed)
    ox, oy = track
    r = f.run(ox[:T], oy[:T], scheme, 0.5) if net is None else \
        f.run_learned(net, ox[:T], oy[:T], scheme, 0.5)
    st = f.state().cpu().numpy()
    W = f.weights().cpu().numpy()
    f.close()
    return r, st, W


def _check(b, tracks, scheme, net=None):
    h = b.histories()
    for f in range(len(NS)):
        print(f"filter {f} n={NS[f]} T={TS[f]}", flush=True)
        r, st, W = _single(NS[f], TS[f], SEEDS[f], tracks[f], scheme, net)
        T = TS[f]
        np.testing.assert_array_equal(h["logz"][f, :T], r["logz"])
        np.testing.assert_array_equal(h["ess"][f, :T], r["ess"])
        np.testing.assert_array_equal(h["resampled"][f, :T], r["resampled"])
        bst, bW = b.filter_state(f)
        np.testing.assert_array_equal(bst.cpu().numpy(), st)
        np.testing.assert_array_equal(bW.cpu().numpy(), W)


@pytest.mark.parametrize("scheme", ["stratified", "systematic"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_filters(dev, scheme, graph):
    import vsmc_amd as va
    sch = va.SCHEMES[scheme]
    tracks = _tracks()
    b = va.CvBatch(NS, TS, SEEDS, sch, 0.5)
    b.set_obs(tracks)
    b.run(graph=graph)
    info = b.info()
    print(info, flush=True)
    assert info["move_blocks"] > 0
    _check(b, tracks, sch)
    # a second run replays from the start with identical results
    h1 = b.histories()
    b.run(graph=graph)
    h2 = b.histories()
    for k in h1:
        np.testing.assert_array_equal(h1[k], h2[k])
    b.close()


def test_batch_learned(dev):
    """Learned proposal in the batched move: identical to the single-filter
    learned runs."""
    import vsmc_amd as va
    rs = np.random.default_rng(5)
    net = va.ProposalNet()
    p = rs.normal(0, 0.1, net.NPARAMS).astype(np.float32)
    net.set_params(p, np.zeros(6, np.float32), np.full(6, 5.0, np.float32))
    tracks = _tracks()
    b = va.CvBatch(NS, TS, SEEDS, va.SYSTEMATIC, 0.5)
    b.set_obs(tracks)
    b.run(net=net, graph=True)
    _check(b, tracks, va.SYSTEMATIC, net)
    b.close()
    net.close()


def test_batch_rejects_unsupported(dev):
    import vsmc_amd as va
    with pytest.raises(va.VsmchError):
        va.CvBatch([10, 20], [5, 5], [1, 2], va.MULTINOMIAL, 0.5)
    with pytest.raises(va.VsmchError):
        va.CvBatch([1, 20], [5, 5], [1, 2], va.SYSTEMATIC, 0.5)
