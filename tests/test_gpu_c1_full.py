"""The GPU particle filter at its full size: N = 2^24 particles, FP32 state,
FP64 weights, systematic resampling when ESS < N/2, 12 steps through the
step-level API (Sampler::do_init / do_iter order, sampler.hpp:735-785). At
every step:

* ESS recomputed in FP64 from the device's normalised weights equals the
  device ESS, and the resampled flag is ESS < N/2 (particle.hpp:172-187);
* at each event: sum of counts = N; counts bit-exact with the oracle's
  systematic scheme (transform.hpp:42-71, u01_sequence.hpp:225-259) fed the
  device's own weights and resampling-stream position; ancestors equal the
  oracle's rep -> index layout (transform.hpp:116-133); the state after the
  event is the pre-event state gathered through them;
* the move of 4,096 sampled particles against the model equations
  (pf_seq.cpp:111-122) in FP64 from the particles' own Philox words.

Then the same 12 steps through vsmch_cv_iterate give the same state,
weights and event count, bit for bit."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
N = 1 << 24
T = 12
SEED = 4242


def c0_obs():
    with open(os.path.join(GOLD, "cv_obs_c0.json")) as f:
        g = json.load(f)
    return (np.array([float.fromhex(x) for x in g["obs_x"]]),
            np.array([float.fromhex(x) for x in g["obs_y"]]))


def _normals_f32(words):
    u = (words >> 8).astype(np.float64) * 2.0 ** -24
    r0, a0 = np.sqrt(-2 * np.log(u[0])), 2 * np.pi * u[1]
    r1, a1 = np.sqrt(-2 * np.log(u[2])), 2 * np.pi * u[3]
    return r0 * np.cos(a0), r0 * np.sin(a0), r1 * np.cos(a1), r1 * np.sin(a1)


def test_c1_full_size_steps(dev):
    import vsmc_amd as va
    ox, oy = c0_obs()
    f = va.CvFilter(N, "float32", seed=SEED)
    k0, kr, _ = f.keys()
    rs = np.random.default_rng(1)
    sample = np.sort(rs.choice(N, 4096, replace=False))
    skeys = (k0 - 1 + sample) % 0xFFFFFFFF + 1
    nres = 0
    for t in range(T):
        if t == 0:
            f.init(float(ox[0]), float(oy[0]))
        else:
            prev = f.state().cpu().numpy()[:, sample].astype(np.float64)
            f.move(t, float(ox[t]), float(oy[t]))
            got = f.state().cpu().numpy()[:, sample].astype(np.float64)
            w = np.array([ol.philox_words(int(k), 4 * t, 4) for k in skeys],
                         dtype=np.uint64).T
            z = _normals_f32(w)
            exp = np.stack([prev[0] + (0.02 * z[0] + 0.1 * prev[2]),
                            prev[1] + (0.02 * z[1] + 0.1 * prev[3]),
                            prev[2] + 0.02 * z[2], prev[3] + 0.02 * z[3]])
            fin = np.isfinite(exp).all(0)
            np.testing.assert_allclose(got[:, fin], exp[:, fin], rtol=2e-6, atol=2e-6)
        W = f.weights().cpu().numpy()
        ess_host = 1.0 / np.dot(W, W)
        ess_dev = f.ess()
        np.testing.assert_allclose(ess_dev, ess_host, rtol=1e-9)
        assert 1.0 <= ess_dev <= N
        _, _, word = f.keys()
        st_before = f.state().cpu().numpy()
        did, rep, idx = f.resample(va.SYSTEMATIC, 0.5, want_rep=True)
        print(f"t={t} ess={ess_dev:.1f} resampled={did}", flush=True)
        assert did == (ess_host < 0.5 * N)
        if did:
            nres += 1
            rep = rep.cpu().numpy()
            idx = idx.cpu().numpy()
            assert rep.sum() == N and (rep >= 0).all()
            erep, used = ol.resample(ol.SYSTEMATIC, W, N, kr, skip=word)
            bad = np.nonzero(rep != erep)[0]
            assert len(bad) == 0, (t, len(bad), bad[:5])
            assert (idx == ol.trans_rep_index(rep)).all()
            st_after = f.state().cpu().numpy()
            assert (st_after == st_before[:, idx]).all()
            assert f.keys()[2] == word + used
            np.testing.assert_allclose(f.ess(), N, rtol=1e-12)
    st, W, nres_dev = f.state().cpu().numpy(), f.weights().cpu().numpy(), f.resample_count()
    f.close()
    assert nres_dev == nres and 0 < nres < T
    # the same steps through Sampler::iterate on the device
    g = va.CvFilter(N, "float32", seed=SEED)
    h = g.iterate(0, ox[:T], oy[:T], va.SYSTEMATIC, 0.5)
    assert (g.state().cpu().numpy() == st).all()
    assert (g.weights().cpu().numpy() == W).all()
    assert g.resample_count() == nres
    assert int(h["resampled"].sum().item()) == nres
    g.close()
