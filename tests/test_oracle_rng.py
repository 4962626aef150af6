"""The oracle's RNG restatement against the reference's own C engines
(include/vsmc/rngc, compiled unchanged into oracle/_ref) and the committed
fixtures those engines produced (tests/golden/rng_words.json, u01.json)."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_philox_stream_golden():
    g = load("rng_words.json")
    for k, v in g["philox"].items():
        k = int(k)
        assert (ol.philox_words(k, 0, 256) == np.array(v["first256"], np.uint32)).all()
        assert (ol.philox_words(k, 1001, 64) == np.array(v["skip1001_64"], np.uint32)).all()


def test_threefry_stream_golden():
    g = load("rng_words.json")
    for k, v in g["threefry"].items():
        k = int(k)
        exp = np.array([int(x) for x in v["first256"]], np.uint64)
        assert (ol.threefry_words(k, 0, 256) == exp).all()
        exp = np.array([int(x) for x in v["skip1001_64"]], np.uint64)
        assert (ol.threefry_words(k, 1001, 64) == exp).all()


def test_blocks_golden():
    g = load("rng_words.json")
    o = ol.oracle()
    for b in g["philox_blocks"]:
        out = np.zeros(4, np.uint32)
        o.vo_philox4x32_10(np.array(b["ctr"], np.uint32), np.array(b["key"], np.uint32), out)
        assert out.tolist() == b["out"]
    for b in g["threefry_blocks"]:
        out = np.zeros(4, np.uint64)
        o.vo_threefry4x64_20(np.array([int(x) for x in b["ctr"]], np.uint64),
                             np.array([int(x) for x in b["key"]], np.uint64), out)
        assert [int(x) for x in out] == [int(x) for x in b["out"]]


def test_u01_golden():
    g = load("u01.json")
    o = ol.oracle()
    for w, d, f in zip(g["w32"], g["u32d"], g["u32f"]):
        assert o.vo_u01_co_u32d(w) == float.fromhex(d)
        assert np.float32(o.vo_u01_co_u32f(w)) == np.float32(float.fromhex(f))
    for w, d in zip(g["w64"], g["u64d"]):
        assert o.vo_u01_co_u64d(int(w)) == float.fromhex(d)
    # closed forms: [0, 1), 32-bit / 53-bit resolution
    assert o.vo_u01_co_u32d(0xFFFFFFFF) == 1 - 2.0 ** -32
    assert o.vo_u01_co_u64d((1 << 64) - 1) == 1 - 2.0 ** -53
    assert o.vo_u01_co_u32f(0xFFFFFFFF) == np.float32(1 - 2.0 ** -24)


@pytest.mark.skipif(not ol.ref_available(), reason="reference tree not built here")
@pytest.mark.parametrize("key", [0, 1, 101, 102, 0xFFFFFFFF, 123456789])
def test_streams_vs_reference_engine(key):
    r = ol.ref()
    for off in (0, 1, 3, 4, 5, 4099):
        a = np.zeros(2000, np.uint32)
        r.ref_philox4x32_words(key, off, 2000, a)
        assert (ol.philox_words(key, off, 2000) == a).all()
        b = np.zeros(2000, np.uint64)
        r.ref_threefry4x64_words(key, off, 2000, b)
        assert (ol.threefry_words(key, off, 2000) == b).all()


def test_engine_loop_equals_batch():
    """CounterEngine scalar stream == batch operator()(n, r) == discard
    (example/rng/include/rng_test.hpp:80-96)."""
    o = ol.oracle()
    rs = np.random.default_rng(3)
    for _ in range(10):
        n = int(rs.integers(1, 5000))
        pre = int(rs.integers(0, 7))
        e1, e2 = ol.Philox(), ol.Philox()
        o.vo_philox_seed(ctypes.byref(e1), 42)
        o.vo_philox_seed(ctypes.byref(e2), 42)
        for _ in range(pre):
            o.vo_philox_next(ctypes.byref(e1))
            o.vo_philox_next(ctypes.byref(e2))
        loop = np.array([o.vo_philox_next(ctypes.byref(e1)) for _ in range(n)], np.uint32)
        batch = np.zeros(n, np.uint32)
        o.vo_philox_fill(ctypes.byref(e2), n, batch)
        assert (loop == batch).all()
        assert o.vo_philox_next(ctypes.byref(e1)) == o.vo_philox_next(ctypes.byref(e2))
        t1, t2 = ol.Threefry(), ol.Threefry()
        o.vo_threefry_seed(ctypes.byref(t1), 7)
        o.vo_threefry_seed(ctypes.byref(t2), 7)
        loop = np.array([o.vo_threefry_next(ctypes.byref(t1)) for _ in range(n)], np.uint64)
        batch = np.zeros(n, np.uint64)
        o.vo_threefry_fill(ctypes.byref(t2), n, batch)
        assert (loop == batch).all()
    # discard(k) == k draws
    for k in (0, 1, 3, 4, 5, 17, 1000):
        e1, e2 = ol.Philox(), ol.Philox()
        o.vo_philox_seed(ctypes.byref(e1), 5)
        o.vo_philox_seed(ctypes.byref(e2), 5)
        for _ in range(k):
            o.vo_philox_next(ctypes.byref(e1))
        o.vo_philox_discard(ctypes.byref(e2), k)
        assert o.vo_philox_next(ctypes.byref(e1)) == o.vo_philox_next(ctypes.byref(e2))


def test_counter_carry():
    """The engine's counter carries into the next word (counter.hpp:42-72):
    after block {2^32-1, 0, 0, 0} comes {0, 1, 0, 0}."""
    o = ol.oracle()
    e = ol.Philox()
    o.vo_philox_seed(ctypes.byref(e), 9)
    o.vo_philox_set_ctr(ctypes.byref(e), np.array([0xFFFFFFFE, 0, 0, 0], np.uint32))
    got = [o.vo_philox_next(ctypes.byref(e)) for _ in range(8)]
    exp = []
    for ctr in ([0xFFFFFFFF, 0, 0, 0], [0, 1, 0, 0]):
        out = np.zeros(4, np.uint32)
        o.vo_philox4x32_10(np.array(ctr, np.uint32), np.array([9, 0], np.uint32), out)
        exp += out.tolist()
    assert got == exp


def test_normals_closed_form():
    """Box-Muller from the pinned U01 values: scalar form pairs (r cos, r sin),
    batch form radii/angles halves (normal_distribution.hpp:199-267)."""
    n = 2049
    u = ol.u01_fill(0, 77, 5, n + 2)
    z = ol.normal_fill(0, 77, 5, n, 0.0, 1.0, form=1)
    r = np.sqrt(-2 * np.log(u[0:n:2]))
    a = 2 * np.pi * u[1:n + 1:2]
    np.testing.assert_allclose(z[0::2], r * np.cos(a), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(z[1::2], (r * np.sin(a))[: len(z[1::2])], rtol=1e-13,
                               atol=1e-13)
    zb = ol.normal_fill(0, 77, 5, n, 1.5, 2.0, form=0)
    for c0 in range(0, n, 1024):
        k = min(1024, n - c0)
        nu = k // 2
        rr = 2.0 * np.sqrt(-2 * np.log(u[c0:c0 + nu]))
        aa = 2 * np.pi * u[c0 + nu:c0 + 2 * nu]
        np.testing.assert_allclose(zb[c0:c0 + nu], rr * np.sin(aa) + 1.5, rtol=1e-12,
                                   atol=1e-12)
        np.testing.assert_allclose(zb[c0 + nu:c0 + 2 * nu], rr * np.cos(aa) + 1.5,
                                   rtol=1e-12, atol=1e-12)
    tail = 1.5 + 2.0 * np.sqrt(-2 * np.log(u[n])) * np.cos(2 * np.pi * u[n + 1])
    assert abs(zb[n - 1] - tail) < 1e-12


def test_normals_distribution():
    from scipy import stats
    z = ol.normal_fill(0, 123, 0, 200000, 0.0, 1.0, form=0)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-4
    u = ol.u01_fill(1, 5, 0, 200000)
    assert stats.kstest(u, "uniform").pvalue > 1e-4
