"""The exact fast-forward of the fp32 pull's wavefront chains (og_chain_fast
in og_pr.hip), restated on the CPU and checked against brute-force
sequential fp32 sums -- the reference's running `sum += contribution`
(pageRank.c:998-1011).

Identity: while the running sum y stays in one binade [2^(e-1), 2^e), the
representable values there are the multiples of u = 2^(e-24), so for a term
a >= 0, fl(y + a) = y + u*m with m the integer nearest a/u, unless a/u ends
in exactly one half (round-half-even then depends on y).  A group of terms
therefore advances y by u * (sum of the m) in one step, provided no term is
a tie and the last sum stays below 2^e; otherwise the group is added term by
term.  The kernel applies this to 512-term groups; the checks below use the
same group size, include tie-heavy sequences (few-bit terms), starts at 0
and at a sum in mid-binade, and a planted mistake (ties not excluded) that
must be caught.
"""
import zlib

import numpy as np
import pytest

G = 512


def _seq(y, a):
    y = np.float32(y)
    for v in a:
        y = np.float32(y + v)
    return y


def _group_fast(y, g, check_ties=True):
    """y advanced over group g by the identity, or None when it does not apply."""
    if not (2.0 ** -100 <= y < 2.0 ** 100):
        return None
    _, e = np.frexp(np.float32(y))
    sc = 24 - int(e)
    t = np.ldexp(g.astype(np.float32), sc).astype(np.float32)
    fl = np.floor(t)
    fr = (t - fl).astype(np.float32)
    bad = np.any(~((t >= 0) & (t < 2.0 ** 24)))
    if check_ties:
        bad = bad or np.any(fr == 0.5)
    if bad:
        return None
    m = fl.astype(np.int64) + (fr > 0.5)
    Y = int(np.ldexp(np.float32(y), sc))
    if Y + int(m.sum()) >= 1 << 24:
        return None
    return np.float32(np.ldexp(np.float32(Y + int(m.sum())), -sc))


def _chain(y, a, check_ties=True):
    y = np.float32(y)
    fast = 0
    for b in range(0, len(a), G):
        g = a[b:b + G]
        f = _group_fast(y, g, check_ties)
        if f is None:
            y = _seq(y, g)
        else:
            y, fast = f, fast + 1
    return y, fast


def _terms(kind, n, rng):
    if kind == "uniform":
        return (rng.random(n) * 1e-3).astype(np.float32)
    if kind == "rank_over_degree":
        return (np.float32(0.15) / rng.integers(1, 64, n).astype(np.float32)).astype(np.float32)
    if kind == "few_bit":              # many exact half-ulp ties
        return np.ldexp(rng.integers(0, 8, n).astype(np.float32), -rng.integers(0, 30, n)).astype(np.float32)
    return (rng.pareto(1.5, n) * 1e-6).astype(np.float32)


@pytest.mark.parametrize("kind", ["uniform", "rank_over_degree", "few_bit", "power_law"])
@pytest.mark.parametrize("y0", [0.0, 0.3, 1e-5])
def test_fast_forward_equals_sequential_sum(kind, y0):
    rng = np.random.default_rng(zlib.crc32(f"{kind}/{y0}".encode()))
    for n in (1, 17, 511, 512, 513, 5000, 20000):
        a = _terms(kind, n, rng)
        want = _seq(y0, a)
        got, _ = _chain(y0, a)
        assert got.view(np.uint32) == want.view(np.uint32), (kind, y0, n)


def test_long_rows_mostly_take_the_fast_path():
    rng = np.random.default_rng(5)
    a = _terms("rank_over_degree", 100_000, rng) * np.float32(1e-3)
    got, fast = _chain(0.0, a)
    assert got.view(np.uint32) == _seq(0.0, a).view(np.uint32)
    assert fast >= 0.9 * (len(a) // G)


def test_planted_tie_mistake_is_caught():
    """Without the tie exclusion the identity rounds half-way terms the
    wrong way on some sums: the brute-force comparison must see it."""
    rng = np.random.default_rng(11)
    wrong = 0
    for _ in range(20):
        a = _terms("few_bit", 4096, rng)
        got, _ = _chain(0.3, a, check_ties=False)
        wrong += got.view(np.uint32) != _seq(0.3, a).view(np.uint32)
    assert wrong > 0
