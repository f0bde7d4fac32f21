"""The exact fast-forward scan of the binned pull (og_bin_wave_sum and
og_bin_long_kernel in og_bin.hip), restated on the CPU and checked against
brute-force sequential fp32 sums -- the reference's running
`sum += contribution` over a row's in-edges (pageRank.c:998-1011).

The scan takes W terms at a time (W = 64 for a wavefront, 1024 for a
block).  With y in the binade of unit u = 2^-sc, term a contributes
m = round(a/u) unless a/u is a tie, negative, not finite or at least 2^24;
an integer prefix sum of the m gives every running sum y + u * prefix up to
the first term that is such a term or whose prefix reaches the binade's top
(Y + prefix >= 2^24, Y = y/u).  That term is added with an ordinary fp32
add, and the scan restarts after it.  A zero or tiny running sum takes one
ordinary add.  Checked on tie-heavy few-bit terms, rank / degree terms,
power-law terms and a run of equal terms; two planted mistakes (ties not
excluded, the binade's top not checked) must be caught.
"""
import zlib

import numpy as np
import pytest


def _seq(a):
    y = np.float32(0.0)
    for v in a:
        y = np.float32(y + v)
    return y


def _scan(a, W, check_ties=True, check_top=True):
    a = np.asarray(a, np.float32)
    y = np.float32(0.0)
    i, n = 0, len(a)
    while i < n:
        if not (2.0 ** -100 <= y < 2.0 ** 100):
            y = np.float32(y + a[i])
            i += 1
            continue
        g = a[i:i + W]
        _, e = np.frexp(y)
        sc = 24 - int(e)
        t = np.ldexp(g, sc).astype(np.float32)
        fl = np.floor(t)
        fr = (t - fl).astype(np.float32)
        bad = ~((t >= 0) & (t < 2.0 ** 24))
        if check_ties:
            bad |= fr == 0.5
        m = np.where(bad, 0, fl.astype(np.int64) + (fr > 0.5))
        incl = np.cumsum(m)
        Y = int(np.ldexp(y, sc))
        fail = bad.copy()
        if check_top:
            fail |= Y + incl >= 1 << 24
        if not fail.any():
            y = np.float32(np.ldexp(np.float32(Y + int(incl[-1])), -sc))
            i += len(g)
        else:
            k = int(np.argmax(fail))
            pre = int(incl[k - 1]) if k else 0
            y = np.float32(np.ldexp(np.float32(Y + pre), -sc))
            y = np.float32(y + g[k])
            i += k + 1
    return y


def _terms(kind, n, rng):
    if kind == "rank_over_degree":
        return (np.float32(0.15) / rng.integers(1, 64, n).astype(np.float32)).astype(np.float32)
    if kind == "few_bit":              # many exact half-ulp ties
        return np.ldexp(rng.integers(0, 8, n).astype(np.float32), -rng.integers(0, 30, n)).astype(np.float32)
    if kind == "equal":
        return np.full(n, np.float32(0.15) / np.float32(3.0), np.float32)
    return (rng.pareto(1.5, n) * 1e-6).astype(np.float32)


@pytest.mark.parametrize("W", [64, 1024])
@pytest.mark.parametrize("kind", ["rank_over_degree", "few_bit", "equal", "power_law"])
def test_scan_equals_sequential_sum(W, kind):
    rng = np.random.default_rng(zlib.crc32(f"{kind}/{W}".encode()))
    for n in (1, 2, 63, 64, 65, 1023, 1025, 5000, 40000):
        a = _terms(kind, n, rng)
        assert _scan(a, W).view(np.uint32) == _seq(a).view(np.uint32), (kind, W, n)


def test_zeros_and_leading_zeros():
    a = np.zeros(300, np.float32)
    assert _scan(a, 64) == 0.0
    a[150:] = np.float32(1e-3)
    assert _scan(a, 64).view(np.uint32) == _seq(a).view(np.uint32)


@pytest.mark.parametrize("mistake", ["ties", "top"])
def test_planted_mistakes_are_caught(mistake):
    """Without the tie exclusion half-way terms round the wrong way; without
    the binade-top check the integer sum runs past 2^e, where the spacing
    doubles: brute-force sums must see both."""
    rng = np.random.default_rng(17)
    wrong = 0
    for _ in range(10):
        a = _terms("few_bit" if mistake == "ties" else "rank_over_degree", 3000, rng)
        got = _scan(a, 64, check_ties=mistake != "ties", check_top=mistake != "top")
        wrong += got.view(np.uint32) != _seq(a).view(np.uint32)
    assert wrong > 0
