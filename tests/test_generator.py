"""Seeded synthetic inputs (host side) -- CPU only.

The generators are this library's own (the reference reads files only), so
they are pinned by an independent Python restatement of their definition
(og_common.h), by the distributions they must follow and by bijectivity of
the vertex permutation.  The device generator is compared with the host one
in the GPU tests.
"""
import numpy as np
import pytest

import opengraph_amd as og

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def stream_key(seed, stream):
    return mix64((seed * GAMMA + stream) & M64)


def draw(key, ctr):
    return mix64((key + (ctr + 1) * GAMMA) & M64)


def perm(seed, bits, x):
    key = stream_key(seed, 1)
    ks = [draw(key, r) for r in range(4)]
    half = (bits + 1) // 2
    hm = (1 << half) - 1
    v = x
    while True:
        L, R = v >> half, v & hm
        for k in ks:
            L, R = R, L ^ (mix64(R ^ k) & hm)
        v = (L << half) | R
        if v < (1 << bits):
            return v


def thr(p):
    x = p * 4294967296.0
    return 0xFFFFFFFF if x >= 4294967295.0 else int(x)


def rmat_edge(scale, a, b, c, seed, permute, e):
    key = stream_key(seed, 0)
    ta, tab, tabc = thr(a), thr(a + b), thr(a + b + c)
    s = d = 0
    for l in range(scale):
        u = draw(key, e * 64 + l) >> 32
        sb = int(u >= tab)
        db = int((ta <= u < tab) or u >= tabc)
        s, d = (s << 1) | sb, (d << 1) | db
    if permute:
        s, d = perm(seed, scale, s), perm(seed, scale, d)
    return s, d


def test_splitmix64_reference_value():
    # SplitMix64 with state 0: first output 0xE220A8397B1DCDAF (published test value)
    assert mix64(GAMMA) == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("scale,permute", [(5, False), (9, True), (16, True)])
def test_rmat_host_matches_restatement(scale, permute):
    n = 300
    src, dst = og.rmat_host(scale, n, 0.57, 0.19, 0.19, 27, permute)
    for e in range(0, n, 7):
        assert (int(src[e]), int(dst[e])) == rmat_edge(scale, 0.57, 0.19, 0.19, 27, permute, e)


def test_rmat_deterministic_and_seeded():
    a = og.rmat_host(12, 5000, seed=3)
    b = og.rmat_host(12, 5000, seed=3)
    c = og.rmat_host(12, 5000, seed=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def test_rmat_quadrant_frequencies():
    """Without permutation the top bits of (src, dst) follow (a, b, c, d)."""
    scale, n = 10, 200000
    src, dst = og.rmat_host(scale, n, 0.57, 0.19, 0.19, 11, False)
    sb, db = src >> (scale - 1), dst >> (scale - 1)
    freq = [np.mean((sb == i) & (db == j)) for i, j in [(0, 0), (0, 1), (1, 0), (1, 1)]]
    assert np.allclose(freq, [0.57, 0.19, 0.19, 0.05], atol=0.005)
    assert src.max() < (1 << scale) and dst.max() < (1 << scale)


@pytest.mark.parametrize("bits", [1, 2, 3, 7, 8, 11, 12])
def test_vertex_permutation_is_a_bijection(bits):
    xs = [perm(5, bits, x) for x in range(1 << bits)]
    assert sorted(xs) == list(range(1 << bits))


def test_permuted_rmat_degree_multiset_unchanged():
    """Permutation relabels vertices: the degree multiset is preserved."""
    s0, d0 = og.rmat_host(11, 40000, seed=8, permute=False)
    s1, d1 = og.rmat_host(11, 40000, seed=8, permute=True)
    assert np.array_equal(np.sort(np.bincount(s0, minlength=2048)), np.sort(np.bincount(s1, minlength=2048)))
    assert np.array_equal(np.sort(np.bincount(d0, minlength=2048)), np.sort(np.bincount(d1, minlength=2048)))


def test_er_uniform():
    src, dst = og.er_host(1000, 400000, 2)
    assert src.max() < 1000 and dst.max() < 1000
    counts = np.bincount(src, minlength=1000)
    assert abs(counts.mean() - 400) < 1e-9 and counts.std() < 3 * np.sqrt(400)


def test_write_bin_layout(tmp_path):
    src = np.array([1, 2, 3], np.uint32)
    dst = np.array([4, 5, 6], np.uint32)
    p = tmp_path / "g.bin"
    og.write_bin(str(p), src, dst)
    assert np.array_equal(np.fromfile(p, np.uint32), [1, 4, 2, 5, 3, 6])
