"""The table of the weights' exponential (vsmc_amd/csrc/common.hpp,
kWexpTab): entry j is 2^(j/64) rounded to double (hi) and the rounded
remainder 2^(j/64) - hi (lo), recomputed here to 90 digits. A wrong entry
would shift every weight whose exponent falls in that 64th of an octave."""
import os
import re
from decimal import Decimal, getcontext

SRC = os.path.join(os.path.dirname(__file__), "..", "vsmc_amd", "csrc", "common.hpp")


def _table():
    text = open(SRC).read()
    body = re.search(r"kWexpTab\[128\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    vals = [float.fromhex(v.strip()) for v in body.replace("\n", " ").split(",") if v.strip()]
    assert len(vals) == 128
    return vals


def test_wexp_table_hi_lo():
    getcontext().prec = 90
    ln2 = Decimal(2).ln()
    tab = _table()
    for j in range(64):
        exact = (Decimal(j) / 64 * ln2).exp()
        hi = float(exact)                       # correctly rounded
        lo = float(exact - Decimal(hi))
        assert tab[2 * j] == hi, (j, tab[2 * j].hex(), hi.hex())
        assert tab[2 * j + 1] == lo, (j, tab[2 * j + 1].hex(), lo.hex())
        # hi + lo carries about 106 bits of 2^(j/64)
        err = abs(Decimal(hi) + Decimal(lo) - exact) / exact
        assert err < Decimal(2) ** -104


def test_wexp_table_monotone():
    tab = _table()
    his = tab[0::2]
    assert all(a < b for a, b in zip(his, his[1:]))
    assert his[0] == 1.0 and his[-1] < 2.0
