"""Pins the oracle's u01 sequence transforms (oracle.c vo_u01_sorted /
vo_u01_stratified / vo_u01_systematic, include/vsmc/rng/u01_sequence.hpp:
54-89) by what they must satisfy: each stratified value lies in its own
stratum [i/n, (i+1)/n); the systematic values start at u/n and step by 1/n;
the sorted values are order statistics (non-decreasing in [0, 1], the n = 1
and n = 2 closed forms 1 - u and 1 - sqrt(u0), 1 - sqrt(u0) u1, and the
Beta(i + 1, n - i) means over repeats)."""
import math

import numpy as np

import oracle_lib as o


def _seq(name, n, arg):
    out = np.zeros(max(n, 1))
    if name == "systematic":
        o.oracle().vo_u01_systematic(n, float(arg), out)
    else:
        getattr(o.oracle(), "vo_u01_" + name)(n, np.ascontiguousarray(arg, np.float64), out)
    return out[:n]


def test_stratified_in_strata():
    rs = np.random.default_rng(1)
    for n in (1, 2, 7, 1000):
        u = rs.random(n)
        u[0] = 0.0
        out = _seq("stratified", n, u)
        i = np.arange(n)
        assert (out >= i / n * (1 - 1e-15)).all() and (out < (i + 1) / n).all()
        np.testing.assert_allclose(out, (u + i) / n, rtol=1e-15, atol=0)


def test_systematic_steps():
    for n in (1, 3, 1000, 4097):
        for u in (0.0, 0.3, 0.999999):
            out = _seq("systematic", n, u)
            assert out[0] == u * (1.0 / n)
            d = np.diff(out)
            np.testing.assert_allclose(d, 1.0 / n, rtol=1e-9)
            np.testing.assert_allclose(out, (u + np.arange(n)) / n, rtol=0, atol=n * 2e-16)


def test_sorted_closed_forms():
    rs = np.random.default_rng(2)
    for _ in range(20):
        u = rs.random(2)
        assert _seq("sorted", 1, u[:1])[0] == 1 - u[0]
        out = _seq("sorted", 2, u)
        assert out[0] == 1 - math.exp(math.log(u[0]) / 2)
        assert math.isclose(out[1], 1 - math.sqrt(u[0]) * u[1], rel_tol=1e-14)


def test_sorted_order_statistics():
    rs = np.random.default_rng(3)
    n, reps = 50, 4000
    acc = np.zeros(n)
    for _ in range(reps):
        out = _seq("sorted", n, rs.random(n))
        assert (np.diff(out) >= 0).all() and out[0] >= 0 and out[-1] <= 1
        acc += out
    mean = acc / reps
    i = np.arange(n)
    exp_mean = (i + 1) / (n + 1)
    sd = np.sqrt((i + 1) * (n - i) / ((n + 1) ** 2 * (n + 2)) / reps)
    assert (np.abs(mean - exp_mean) < 5 * sd).all()
