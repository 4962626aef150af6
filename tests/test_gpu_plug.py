"""The reference's plug points over the HIP layer (tests/cpp/plug_hip.cpp,
whose static_asserts check the member set at compile time): an event driven
by Particle::resample's control flow (particle.hpp:170-197) through
StateHIP::copy(N, nullptr), StateHIP::copy with a host index array, and the
Weight members set / mul / set_log / add_log (plain and strided), data,
read_weight and draw, all against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "plug_hip")


def seed_key(s, i):
    m = 0xFFFFFFFF
    return (s % m + i) % m + 1


def _norm(w):
    return w / w.sum()


@pytest.mark.parametrize("scheme", [ol.STRATIFIED, ol.SYSTEMATIC])
def test_plug_points(dev, tmp_path, scheme):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    N = 20011
    rs = np.random.default_rng(scheme)
    x = rs.normal(size=(3, N))
    lw = rs.normal(0, 2, N)
    W = np.exp(lw - lw.max())
    W /= W.sum()
    idx = ol.trans_rep_index(ol.trans_u01_rep(W, np.sort(rs.random(N))))
    w2 = rs.uniform(0.1, 2.0, 2 * N)
    inp = tmp_path / "in.txt"
    with open(inp, "w") as f:
        f.write(" ".join(v.hex() for v in x.ravel()) + "\n")
        f.write(" ".join(v.hex() for v in lw) + "\n")
        f.write(" ".join(str(int(v)) for v in idx) + "\n")
        f.write(" ".join(v.hex() for v in w2) + "\n")
    r = subprocess.run([EXE, str(N), str(scheme), str(inp)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    out, draws = {}, {"DRAW32": [], "DRAW64": []}
    for ln in r.stdout.strip().split("\n"):
        tag, *vals = ln.split("\t")
        if tag in draws:
            draws[tag].append((int(vals[0]), int(vals[1])))
        elif tag in ("SEED", "RESAMPLED"):
            out[tag] = int(vals[0])
        else:
            out[tag] = np.array([float.fromhex(v) for v in vals])
    # set_log from a host array, ESS, the decision
    W0 = out["W0"]
    np.testing.assert_allclose(W0, ol.weight_set_log(lw)[0], rtol=1e-12)
    np.testing.assert_allclose(out["ESS0"][0], 1 / np.dot(W0, W0), rtol=1e-12)
    assert out["RESAMPLED"] == int(out["ESS0"][0] < 0.5 * N) == 1
    # the event through copy(N, nullptr): the oracle's counts from the same
    # weights and the resampling engine seed_key(seed, N), gathered state
    erep, _ = ol.resample(scheme, W0, N, seed_key(out["SEED"], N))
    eidx = ol.trans_rep_index(erep)
    X1 = out["X1"].reshape(3, N)
    assert (X1 == x[:, eidx]).all()
    assert (out["W1"] == 1.0 / N).all()
    # copy with a host index array
    assert (out["X2"].reshape(3, N) == X1[:, idx]).all()
    # strided linear set and mul, log forms
    np.testing.assert_allclose(out["WSET"], _norm(w2[0::2]), rtol=1e-12)
    np.testing.assert_allclose(out["WMUL"], _norm(w2[0::2] * w2[1::2]), rtol=1e-12)
    np.testing.assert_allclose(out["WSETLOG"], W0, rtol=1e-12)
    wadd = _norm(out["WSETLOG"] * np.exp(w2[0::2] - w2[0::2].max()))
    np.testing.assert_allclose(out["WADDLOG"], wadd, rtol=1e-12)
    assert (out["DATA"] == out["WADDLOG"]).all()
    assert (out["READ2"][0::2] == out["WADDLOG"]).all() and (out["READ2"][1::2] == -1).all()
    # draw: the first index whose sequential cumulative weight is >= u
    c = ol.cumsum(out["WADDLOG"])
    for tag, conv in (("DRAW32", lambda w: w * 2.0 ** -32),
                      ("DRAW64", lambda w: (w >> 11) * 2.0 ** -53)):
        assert len(draws[tag]) == 8
        for word, got in draws[tag]:
            u = conv(word)
            k = int(np.searchsorted(c, u, side="left"))
            assert got == min(k, N - 1), (tag, word, got, k)


def test_python_copy_and_strided_updates(dev):
    """The same members through the C ABI from Python."""
    import torch
    import vsmc_amd as va
    n = 5003
    rs = np.random.default_rng(4)
    p = va.Particle(n, 2, "float32", seed=3)
    x = torch.as_tensor(rs.normal(size=(2, n)), dtype=torch.float32, device="cuda")
    p.set_state(x)
    w = rs.uniform(0.5, 1.5, 3 * n)
    p.weight_update(va.WEIGHT_SET, torch.as_tensor(w), stride=3)
    np.testing.assert_allclose(p.weights().cpu().numpy(), _norm(w[0::3]), rtol=1e-12)
    p.weight_update(va.WEIGHT_MUL, torch.as_tensor(w[1:]), stride=3)
    np.testing.assert_allclose(p.weights().cpu().numpy(), _norm(w[0::3] * w[1::3]), rtol=1e-12)
    idx = rs.integers(0, n, n)
    p.copy(torch.as_tensor(idx))
    assert (p.state().cpu().numpy() == x.cpu().numpy()[:, idx]).all()
    with pytest.raises(va.VsmchError):
        p.copy(torch.as_tensor(np.full(n, n)))   # out of range: rejected
    assert (p.state().cpu().numpy() == x.cpu().numpy()[:, idx]).all()
    with pytest.raises(va.VsmchError):
        p.weight_update(va.WEIGHT_SET, torch.as_tensor(w), stride=0)
    p.close()
