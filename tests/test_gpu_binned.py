"""The binned fp32 pull (og_bin.hip, the default for variant 0) against the
oracle's sequential pull and the reference's golden outputs, through the C
ABI.

Bar: bit-exact.  Every row is the reference's running fp32 sum over its
in-edges in CSR order (pageRank.c:998-1011): thread rows add from LDS one
term at a time; wavefront rows and rows longer than a bin (cut into pieces,
added by one block) use the exact binade fast-forward, pinned on the CPU by
tests/test_bin_scan.py.  Cases: every golden fp32-pull case; a graph with
hubs just under, at and well over the bin capacity whose in-neighbours have
small out-degrees (few-bit contributions: many half-ulp ties in the long
sums); R-MAT scales 18 and 20; Erdos-Renyi with a vertex count that is
not a multiple of 4; tiny graphs.  Ranks, iteration counts and
per-iteration activeVertices must all match.
"""
import hashlib

import numpy as np
import pytest

import opengraph_amd as og
from conftest import golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

CAP = 19968   # OG_BIN_CAP: in-edge values per bin


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _run(g, eps, iters, **kw):
    pr = og.PageRank(g, 0, iters, eps, **kw)
    res = pr.run()
    x = pr.ranks()
    act, err = pr.rows()
    pr.free()
    return res, x, act, err


def _check(g, o, eps, iters, label):
    ref = po.pagerank(o, 0, eps, iters, 8)
    res, x, act, err = _run(g, eps, iters, binned=1)
    assert np.array_equal(x.view(np.uint32), ref["ranks"].view(np.uint32)), label
    assert res.iterations == ref["iterations"], label
    assert list(act) == list(ref["active"]), label
    np.testing.assert_allclose(err, ref["errors"], rtol=1e-9, atol=0)
    return res


def test_binned_bitexact_on_golden_cases(golden):
    """Every golden fp32-pull case (graphbrew, LAW-enron, R-MAT-16; sorts 0
    and 1; both tolerances) with the default options."""
    graphs = {}
    n = 0
    for c in golden["cases"]:
        if c["code"] != 0:
            continue
        key = (c["graph"], c["sort"])
        if key not in graphs:
            s, d = golden_input(c["graph"], golden)
            graphs[key] = og.Graph.from_edges(s, d, c["sort"])
        res, x, act, _ = _run(graphs[key], c["epsilon"], c["max_iter"])
        assert sha(x) == c["ranks_sha256"], c
        assert res.iterations == c["iterations"]
        assert list(act) == c["active"]
        assert res.error_total == pytest.approx(c["error_total"], rel=1e-9, abs=1e-300)
        n += 1
    for g in graphs.values():
        g.free()
    assert n >= 10


def _hub_graph(seed):
    """R-MAT-14 background (ids below 2^14), leaves (the next 200,000 ids,
    no in-edges, so their contributions are 0.15f over small out-degrees:
    few-bit values) and four hubs with exactly CAP - 1, CAP, CAP + 1 and
    2.5 * CAP in-edges from the leaves; the leaves also point into the
    background."""
    rng = np.random.default_rng(seed)
    s, d = og.rmat_host(14, 16 << 14, seed=seed)
    src, dst = [s], [d]
    n0 = 1 << 14
    for k, cnt in enumerate((CAP - 1, CAP, CAP + 1, 5 * CAP // 2)):
        leaves = (n0 + rng.integers(0, 200_000, cnt)).astype(np.uint32)
        src.append(leaves)
        dst.append(np.full(cnt, n0 + 200_000 + k, np.uint32))
    leaves = np.unique(np.concatenate(src[1:]))
    extra = rng.integers(0, 4, len(leaves))          # 0, 1, 3 or 7 more out-edges
    for k, more in ((1, 1), (2, 3), (3, 7)):
        sel = np.repeat(leaves[extra == k], more)
        src.append(sel)
        dst.append(rng.integers(0, n0, len(sel)).astype(np.uint32))
    return np.concatenate(src).astype(np.uint32), np.concatenate(dst).astype(np.uint32)


def test_binned_rows_longer_than_a_bin():
    s, d = _hub_graph(3)
    g = og.Graph.from_edges(s, d, 0)
    o = po.build_csr(s, d, 0, 8)
    indeg = o["i_out_degree"]
    assert indeg.max() > 2 * CAP and (indeg == CAP).any() and (indeg == CAP + 1).any()
    _check(g, o, 0.0, 8, "hubs, 8 iterations")
    _check(g, o, 1e-7, 100, "hubs, to 1e-7")
    g.free()


@pytest.mark.parametrize("scale", [18, 20])
def test_binned_bitexact_rmat(scale):
    s, d = og.rmat_host(scale, 16 << scale, seed=scale)
    g = og.Graph.from_edges(s, d, 0)
    o = po.build_csr(s, d, 0, 8)
    print("scale", scale, "max in-degree", int(o["i_out_degree"].max()), flush=True)
    _check(g, o, 0.0, 10, f"rmat{scale}")
    res = _check(g, o, 1e-6, 100, f"rmat{scale} to 1e-6")
    assert res.iterations > 10
    g.free()


def test_binned_er_and_tiny_graphs():
    s, d = og.er_host(100_003, 1_600_000, 9)
    cases = [(s, d)]
    cases.append((np.array([0], np.uint32), np.array([0], np.uint32)))
    cases.append((np.array([3, 1, 3, 0, 2, 2], np.uint32), np.array([1, 3, 0, 3, 2, 2], np.uint32)))
    cases.append((np.array([5, 5, 5, 9], np.uint32), np.array([2, 2, 2, 5], np.uint32)))
    for i, (s, d) in enumerate(cases):
        g = og.Graph.from_edges(s, d, 0)
        o = po.build_csr(s, d, 0, 4)
        _check(g, o, 0.0, 12, f"case {i}")
        _check(g, o, 1e-7, 100, f"case {i} to 1e-7")
        g.free()


def test_binned_option_rules():
    """binned = 1 asks for the binned pull and fails where it does not
    apply; the default quietly keeps the gather pull there."""
    s, d = og.rmat_host(10, 8 << 10, seed=1)
    g = og.Graph.from_edges(s, d, 0)
    with pytest.raises(og.OgpuError, match="binned"):
        og.PageRank(g, 2, 10, 1e-6, binned=1)
    with pytest.raises(og.OgpuError, match="binned"):
        og.PageRank(g, 0, 10, 1e-6, binned=1, ppr_source=3)
    og.PageRank(g, 2, 10, 1e-6).free()
    g.free()
    e = og.Graph.from_edges(np.zeros(0, np.uint32), np.zeros(0, np.uint32), 0)
    with pytest.raises(og.OgpuError, match="binned"):
        og.PageRank(e, 0, 10, 1e-6, binned=1)
    res, x, _, _ = _run(e, 1e-6, 10)
    e.free()
