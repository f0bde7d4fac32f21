"""Pins of the personalized PageRank oracles (CPU only).

ppr_fp64 is pinned by the dense solve of x = (1-d) e_s + d P^T D^-1 x (no
dangling redistribution) and by a planted wrong restart vertex; ppr_f32 (the
GPU kernels' arithmetic) is pinned against ppr_fp64 and by its first
iterations written out by hand.
"""
import numpy as np
import pytest

from conftest import golden_input
from oracle import pyoracle as po


def _dense(g):
    N = g.num_vertices
    P = np.zeros((N, N))
    for u, v in zip(g["f_src"], g["f_dst"]):
        P[u, v] += 1.0 / g["f_out_degree"][u]
    return P


@pytest.fixture(scope="module")
def graphbrew(golden):
    s, d = golden_input("graphbrew", golden)
    return po.build_csr(s, d, 0, 1)


@pytest.mark.parametrize("source", [0, 3, 7])
def test_ppr_fp64_dense_solve(graphbrew, source):
    g = graphbrew
    N = g.num_vertices
    e = np.zeros(N)
    e[source] = 0.15
    want = np.linalg.solve(np.eye(N) - 0.85 * _dense(g).T, e)
    x, it, l1 = po.ppr_fp64(g, source, 1e-15, 5000)
    assert l1 < 1e-15
    np.testing.assert_allclose(x, want, rtol=1e-11, atol=1e-15)


def test_planted_wrong_restart_is_caught(graphbrew):
    g = graphbrew
    N = g.num_vertices
    e = np.zeros(N)
    e[4] = 0.15
    wrong = np.linalg.solve(np.eye(N) - 0.85 * _dense(g).T, e)
    x, _, _ = po.ppr_fp64(g, 3, 1e-15, 5000)
    assert not np.allclose(x, wrong, rtol=1e-3)


@pytest.mark.parametrize("source", [1, 5])
def test_ppr_f32_close_to_fp64(graphbrew, source):
    x32, it32, hist = po.ppr_f32(graphbrew, source, 1e-7, 300)
    x64, _, _ = po.ppr_fp64(graphbrew, source, 1e-15, 5000)
    assert hist[-1] < 1e-7 and all(h >= 1e-7 for h in hist[:-1])
    assert np.abs(x32.astype(np.float64) - x64).sum() < 1e-5


def test_ppr_f32_first_iterations_by_hand(graphbrew):
    """Iteration 1 from e_s: x1 = (1-d) e_s + d * (1/outdeg(s)) on s's
    out-neighbours (with multiplicity), in fp32."""
    g = graphbrew
    s = int(np.flatnonzero(g["f_out_degree"] > 0)[2])
    x1, it, hist = po.ppr_f32(g, s, 0.0, 1)
    assert it == 1
    deg = np.float32(g["f_out_degree"][s])
    c = np.float32(1.0) / deg
    want = np.zeros(g.num_vertices, np.float32)
    nbrs = g["f_dst"][g["f_edges_idx"][s]: g["f_edges_idx"][s] + g["f_out_degree"][s]]
    cnt = np.bincount(nbrs, minlength=g.num_vertices)
    for v in np.flatnonzero(cnt):
        acc = np.float32(0.0)
        for _ in range(cnt[v]):
            acc = np.float32(acc + c)
        want[v] = np.float32(np.float32(0.85) * acc)
    want[s] = np.float32(np.float32(np.float32(1.0) - np.float32(0.85)) + want[s])
    assert np.array_equal(x1, want)
    assert hist[0] == pytest.approx(np.abs(want.astype(np.float64) - (np.arange(g.num_vertices) == s)).sum(),
                                    rel=1e-12)


def test_ppr_dangling_source():
    """A source without out-edges: x = (1-d) e_s after one iteration, L1 = d,
    then L1 = 0."""
    g = po.build_csr(np.array([1, 2], np.uint32), np.array([2, 0], np.uint32), 0, 1)
    x, it, hist = po.ppr_f32(g, 0, 1e-9, 10)
    assert it == 2
    assert x[0] == np.float32(1.0) - np.float32(0.85) and x[1] == 0 and x[2] == 0
    assert hist[0] == pytest.approx(0.85, rel=1e-7) and hist[1] == 0.0
