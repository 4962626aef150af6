"""HIP stream fills against the oracle and the reference-made fixtures:
words and U01 bit-exact; normals within 1e-12 (libm-dependent)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _np(t):
    return t.cpu().numpy()


def test_words_golden(dev):
    import vsmc_amd as va
    with open(os.path.join(GOLD, "rng_words.json")) as f:
        g = json.load(f)
    for k, v in g["philox"].items():
        got = _np(va.rng_words(va.PHILOX4X32, int(k), 0, 256)).view(np.uint32)
        assert (got == np.array(v["first256"], np.uint32)).all(), k
        got = _np(va.rng_words(va.PHILOX4X32, int(k), 1001, 64)).view(np.uint32)
        assert (got == np.array(v["skip1001_64"], np.uint32)).all(), k
    for k, v in g["threefry"].items():
        got = _np(va.rng_words(va.THREEFRY4X64, int(k), 0, 256)).view(np.uint64)
        assert (got == np.array([int(x) for x in v["first256"]], np.uint64)).all(), k


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("key,offset,n", [(0, 0, 1), (1, 3, 7), (101, 5, 4099),
                                          (0xFFFFFFFF, 2, 1 << 20), (102, 1 << 33, 1000)])
def test_words_and_u01_bit_exact(dev, kind, key, offset, n):
    import vsmc_amd as va
    print(f"kind={kind} key={key} offset={offset} n={n}", flush=True)
    w = _np(va.rng_words(kind, key, offset, n))
    if kind == 0:
        assert (w.view(np.uint32) == ol.philox_words(key, offset, n)).all()
    else:
        assert (w.view(np.uint64) == ol.threefry_words(key, offset, n)).all()
    u = _np(va.u01_fill(kind, key, offset, n))
    assert (u == ol.u01_fill(kind, key, offset, n)).all()


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("n", [1, 99, 1023, 1024, 1025, 4097, 100001])
def test_normals(dev, kind, n):
    import vsmc_amd as va
    print(f"kind={kind} n={n}", flush=True)
    for form in (0, 1):
        got = _np(va.normal_fill(kind, 9, 3, n, 0.5, 2.0, form=form))
        exp = ol.normal_fill(kind, 9, 3, n, 0.5, 2.0, form=form)
        np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)


def test_full_size_sampled(dev):
    """2^26 U01 doubles (Philox) and words (Threefry): random slices equal the
    oracle's stream at those offsets; mean and range of the whole output."""
    import torch
    import vsmc_amd as va
    n = 1 << 26
    u = va.u01_fill(0, 2024, 0, n)
    torch.cuda.synchronize()
    assert 0.499 < float(u.mean()) < 0.501
    assert float(u.min()) >= 0.0 and float(u.max()) < 1.0
    rs = np.random.default_rng(1)
    for off in rs.integers(0, n - 512, 8):
        off = int(off)
        assert (_np(u[off:off + 512]) == ol.u01_fill(0, 2024, off, 512)).all()
    w = va.rng_words(1, 2025, 0, 1 << 22)
    for off in rs.integers(0, (1 << 22) - 64, 8):
        off = int(off)
        assert (_np(w[off:off + 64]).view(np.uint64) == ol.threefry_words(2025, off, 64)).all()
