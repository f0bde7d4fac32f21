"""Personalized PageRank on the GPU: the single-source personalized option of
the pull kernel and the batched N x k pull-SpMM, through the C ABI.

Bars: single-source in CSR-order mode (hub_threshold 0xFFFFFFFF) is bit-exact
against the oracle's fp32 restatement (ranks and iteration count); the batched
block is bit-exact per column against the single-source run on every row with
at most 512 in-edges and within 1e-5 relative elsewhere (longer rows add
512-edge chunk sums); every column is within 1e-5 (L1) of the FP64 oracle.
"""
import numpy as np
import pytest

import opengraph_amd as og
from conftest import golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ALL_EXACT = 0xFFFFFFFF


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


@pytest.fixture(scope="module")
def rmat16(golden):
    s, d = golden_input("rmat16", golden)
    g = og.Graph.from_edges(s, d, 0)
    o = po.build_csr(s, d, 0, 8)
    yield g, o
    g.free()


def _single(g, src, eps=1e-6, iters=100, hub=ALL_EXACT):
    pr = og.PageRank(g, 0, iters, eps, hub_threshold=hub, ppr_source=int(src))
    res = pr.run()
    x = pr.ranks()
    _, err = pr.rows()
    pr.free()
    return x, res, err


def _sources(n, k, seed):
    return np.random.default_rng(seed).integers(0, n, size=k, dtype=np.uint32)


@pytest.mark.parametrize("src_index", [0, 1, 2])
def test_single_source_bitexact_vs_oracle(rmat16, src_index):
    g, o = rmat16
    src = int(_sources(g.num_vertices, 3, 11)[src_index])
    x, res, err = _single(g, src)
    ref, it, hist = po.ppr_f32(o, src, 1e-6, 100)
    print("source", src, "iterations", res.iterations, it, flush=True)
    assert np.array_equal(x.view(np.uint32), ref.view(np.uint32))
    assert res.rows == it
    np.testing.assert_allclose(err[:it], hist, rtol=1e-9, atol=0)


def test_single_source_default_mode_close_to_fp64(rmat16):
    g, o = rmat16
    src = int(_sources(g.num_vertices, 1, 12)[0])
    x, res, _ = _single(g, src, hub=0)
    x64, _, _ = po.ppr_fp64(o, src, 1e-13, 1000)
    assert np.abs(x.astype(np.float64) - x64).sum() <= 1e-5


def test_single_source_rejects_bad_options(rmat16):
    g, _ = rmat16
    with pytest.raises(og.OgpuError):
        og.PageRank(g, 2, 10, 1e-6, ppr_source=0)          # fixed point: not personalized
    with pytest.raises(og.OgpuError):
        og.PageRank(g, 0, 10, 1e-6, ppr_source=g.num_vertices)


@pytest.fixture(scope="module")
def er16():
    """Uniform graph, in-degrees far below 512: every batched row is summed in
    CSR order, so each column must equal the single-source run bit for bit."""
    s, d = og.er_host(1 << 16, 1 << 20, 21)
    g = og.Graph.from_edges(s, d, 0)
    o = po.build_csr(s, d, 0, 8)
    assert o["i_out_degree"].max() <= 512
    yield g, o
    g.free()


@pytest.mark.parametrize("k", [1, 3, 64])
def test_batched_bitexact_vs_single_source_runs(er16, k):
    g, o = er16
    src = _sources(g.num_vertices, k, 100 + k)
    p = og.Ppr(g, src, 100, 1e-6)
    res = p.run()
    X = p.ranks()
    iters, hist = p.columns()
    p.free()
    assert res.converged == k
    # every column: converged columns move out of the gathered slots during
    # the run, and each must still match its own single-source run
    for j in range(k):
        x, r1, err = _single(g, src[j])
        assert iters[j] == r1.rows, j
        assert np.array_equal(X[:, j].view(np.uint32), x.view(np.uint32)), j
        np.testing.assert_allclose(hist[:iters[j], j], err[:iters[j]], rtol=1e-12, atol=0)


def test_batched_close_to_single_source_on_skewed_graph(rmat16):
    """R-MAT has rows of thousands of in-edges (chunked in the batch)."""
    g, o = rmat16
    assert o["i_out_degree"].max() > 512
    k = 16
    src = _sources(g.num_vertices, k, 5)
    p = og.Ppr(g, src, 100, 1e-6)
    res = p.run()
    X = p.ranks()
    iters, _ = p.columns()
    p.free()
    assert res.converged == k
    for j in range(0, k, 5):
        x, r1, _ = _single(g, src[j])
        assert abs(int(iters[j]) - int(r1.rows)) <= 1, j
        np.testing.assert_allclose(X[:, j], x, rtol=1e-5, atol=1e-9)


def test_batched_vs_fp64_oracle(rmat16):
    g, o = rmat16
    k = 8
    src = _sources(g.num_vertices, k, 7)
    p = og.Ppr(g, src, 100, 1e-6)
    p.run()
    X = p.ranks()
    p.free()
    for j in range(k):
        x64, _, _ = po.ppr_fp64(o, int(src[j]), 1e-13, 1000)
        l1 = np.abs(X[:, j].astype(np.float64) - x64).sum()
        print("column", j, "source", src[j], "L1 vs fp64", l1, flush=True)
        assert l1 <= 1e-5


def test_batched_duplicate_and_dangling_sources(rmat16):
    """Two identical columns stay identical; a source without out-edges keeps
    exactly (1-d) at itself and zero elsewhere after its first iteration."""
    g, o = rmat16
    outdeg = o["f_out_degree"]
    dangling = int(np.flatnonzero(outdeg == 0)[0])
    s = int(np.flatnonzero(outdeg > 0)[5])
    p = og.Ppr(g, [s, s, dangling], 100, 1e-6)
    res = p.run()
    X = p.ranks()
    iters, _ = p.columns()
    p.free()
    assert res.converged == 3
    assert np.array_equal(X[:, 0].view(np.uint32), X[:, 1].view(np.uint32))
    want = np.zeros(g.num_vertices, np.float32)
    want[dangling] = np.float32(1.0) - np.float32(0.85)
    assert np.array_equal(X[:, 2], want)
    assert iters[2] == 2


def test_batched_iteration_cap(rmat16):
    """Sources with out-edges (a source without any converges in two
    iterations): all four columns run into the cap."""
    g, o = rmat16
    src = np.flatnonzero(o["f_out_degree"] > 0)[[3, 50, 500, 5000]].astype(np.uint32)
    p = og.Ppr(g, src, 5, 1e-12)
    res = p.run()
    iters, hist = p.columns()
    p.free()
    assert res.iterations == 5 and res.converged == 0
    assert list(iters) == [5] * 4 and hist.shape == (5, 4)


def test_batched_rejects_bad_arguments(rmat16):
    g, _ = rmat16
    with pytest.raises(og.OgpuError):
        og.Ppr(g, np.zeros(65, np.uint32))
    with pytest.raises(og.OgpuError):
        og.Ppr(g, [g.num_vertices])
