"""SpMV (-p 0..3 of SPMVGraphCSR) on the GPU, through the C ABI.

Bar: bit-exact against the compiled reference's vector_output for every
golden case (graphbrew, LAW-enron, R-MAT-16; sort modes 0 and 1; 20
iterations; tests/golden/make_golden_spmv.py) for fp32 pull with every row
in the reference's order, fixed-point pull, and fixed-point push both
gathered and scattered.  The default fp32 pull (binned) is bit-exact on
every row; with a given row split (the gather pull) rows of at most 16
in-edges stay bit-exact and the longer ones lie within the fp32 summation
bound of the exact value, as the reference's own running sums do; fp32 push
(unordered atomics, as the reference's omp atomic) is held to the same
bound.  Larger graphs (R-MAT-20) against the oracle, itself pinned on
those fixtures; the drop-in SPMVGraphCSR on host GraphCSR structs.
"""
import hashlib

import numpy as np
import pytest

import opengraph_amd as og
from conftest import golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ALL_EXACT = 1 << 31


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


def _exact_and_bound(csr, iterations):
    """fp64 value of the running sum and its fp32 sequential-summation bound."""
    N = csr.num_vertices
    od = csr["f_out_degree"].astype(np.float64)
    x = np.where(od > 0, (1.0 / np.maximum(od, 1)).astype(np.float32), np.float32(0.001)).astype(np.float64)
    src, dst = csr["f_src"].astype(np.int64), csr["f_dst"].astype(np.int64)
    want = iterations * 0.0001 * np.bincount(dst, weights=x[src], minlength=N)
    terms = iterations * np.bincount(dst, minlength=N).astype(np.float64)
    return want, (terms + 2.0) * 2.0 ** -24 * want + 1e-12


def _run(g, code, iterations=20, **kw):
    s = og.Spmv(g, code, **kw)
    try:
        s.run(iterations)
        return s.output()
    finally:
        s.free()


def test_golden_cases_bit_exact(golden):
    graphs = {}
    try:
        for key, c in golden["spmv"].items():
            gname, sort, code = key.split("/")
            sort, code = int(sort[4:]), int(code[1:])
            if (gname, sort) not in graphs:
                s, d = golden_input(gname, golden)
                graphs[(gname, sort)] = og.Graph.from_edges(s, d, sort)
            g = graphs[(gname, sort)]
            it = c["iterations"]
            if code == 0:
                for hot in (0, 1, 2):
                    y = _run(g, 0, it, hub_threshold=ALL_EXACT, hot_cache=hot)
                    assert sha(y) == c["output_sha256"], (key, hot)
            elif code in (2, 3):
                for hot in (0, 1, 2):
                    y = _run(g, code, it, hot_cache=hot)
                    assert sha(y) == c["output_sha256"], (key, hot)
                if code == 3:
                    for hot in (0, 1):
                        y = _run(g, 3, it, scatter=1, hot_cache=hot)
                        assert sha(y) == c["output_sha256"], (key, "scatter", hot)
            print(key, "ok", flush=True)
    finally:
        for g in graphs.values():
            g.free()


@pytest.mark.parametrize("gname", ["enron", "rmat16"])
def test_default_rows_and_push_within_bound(golden, gname):
    s, d = golden_input(gname, golden)
    csr = po.build_csr(s, d, 0, 4)
    ref = po.spmv(csr, 0, 20)
    want, bound = _exact_and_bound(csr, 20)
    short = np.bincount(csr["f_dst"], minlength=csr.num_vertices) <= 16
    g = og.Graph.from_edges(s, d, 0)
    try:
        # the default (binned) pull: every row's running y in the reference's order
        y = _run(g, 0, 20)
        assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
        for hot in (0, 1):     # the gather pull with its row split
            y = _run(g, 0, 20, hub_threshold=16, hot_cache=hot)
            assert np.array_equal(y[short].view(np.uint32), ref[short].view(np.uint32)), hot
            assert np.all(np.abs(y.astype(np.float64) - want) <= bound), hot
        for hot in (0, 2):          # LDS accumulators (default for -p 1) and plain atomics
            y = _run(g, 1, 20, hot_cache=hot)
            assert np.all(np.abs(y.astype(np.float64) - want) <= bound), ("push", hot)
            # the reference's (single-thread) output obeys the same bound
            assert np.all(np.abs(y.astype(np.float64) - ref) <= 2 * bound), ("push", hot)
    finally:
        g.free()


def test_iteration_counts_and_edge_cases(golden):
    s, d = golden_input("graphbrew", golden)
    csr = po.build_csr(s, d, 1, 1)
    g = og.Graph.from_edges(s, d, 1)
    try:
        for code in range(4):
            assert not np.any(_run(g, code, 0)), code          # no pass: y stays 0
            for it in (1, 3):
                y = _run(g, code, it, hub_threshold=ALL_EXACT)
                want = po.spmv(csr, code, it)
                if code == 1:
                    _, bound = _exact_and_bound(csr, it)
                    assert np.all(np.abs(y - want) <= 2 * bound)
                else:
                    assert np.array_equal(y.view(np.uint32), want.view(np.uint32)), (code, it)
        with pytest.raises(og.OgpuError):
            og.Spmv(g, 4)
        sp = og.Spmv(g, 0)
        try:
            sp.run(5)
            first = sp.output()
            sp.run(5)                                        # a second run starts from y = 0
            assert np.array_equal(first, sp.output())
        finally:
            sp.free()
    finally:
        g.free()


@pytest.fixture(scope="module")
def rmat20():
    s, d = og.rmat_host(20, 16 << 20, seed=9)
    g = og.Graph.from_edges(s, d, 0)
    csr = po.build_csr(s, d, 0, 8)
    yield g, csr
    g.free()


def test_scale20_fixed_point_bit_exact(rmat20):
    g, csr = rmat20
    for code in (2, 3):
        want = po.spmv(csr, code, 8)
        assert np.array_equal(_run(g, code, 8).view(np.uint32), want.view(np.uint32)), code
    want = po.spmv(csr, 3, 8)
    assert np.array_equal(_run(g, 3, 8, scatter=1).view(np.uint32), want.view(np.uint32))


def test_scale20_fp32(rmat20):
    g, csr = rmat20
    ref = po.spmv(csr, 0, 8)
    want, bound = _exact_and_bound(csr, 8)
    short = np.bincount(csr["f_dst"], minlength=csr.num_vertices) <= 16
    y = _run(g, 0, 8)                          # binned: bit-exact on every row
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
    y = _run(g, 0, 8, hub_threshold=16)        # gather pull with its row split
    assert np.array_equal(y[short].view(np.uint32), ref[short].view(np.uint32))
    assert np.all(np.abs(y.astype(np.float64) - want) <= bound)
    y = _run(g, 0, 8, hub_threshold=ALL_EXACT)
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
    y = _run(g, 1, 8)
    assert np.all(np.abs(y.astype(np.float64) - want) <= bound)


def test_dropin_spmv_graph_csr(golden):
    for gname in ("graphbrew", "rmat16"):
        s, d = golden_input(gname, golden)
        csr = po.build_csr(s, d, 0, 4)
        host = og.HostGraphCSR(csr.a)
        od = csr["f_out_degree"]
        x_want = np.where(od > 0, np.float32(1.0) / np.maximum(od, 1).astype(np.float32), np.float32(0.001))
        for code in (0, 2, 3, 7):
            y, x, it, t = og.spmv_graph_csr(host, code, 20)
            assert it == 20 and t > 0.0
            assert np.array_equal(x.view(np.uint32), x_want.astype(np.float32).view(np.uint32))
            want = po.spmv(csr, code if code <= 3 else 0, 20)       # SPMV.c:609-611: default pull
            # fixed point, and the fp32 pull (binned by default): every row bit-exact
            assert np.array_equal(y.view(np.uint32), want.view(np.uint32)), (gname, code)
        y, _, _, _ = og.spmv_graph_csr(host, 1, 20)
        want, bound = _exact_and_bound(csr, 20)
        assert np.all(np.abs(y.astype(np.float64) - want) <= bound)
