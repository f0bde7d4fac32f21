"""GPU parity at the full workloads the benchmarks are quoted on, through the
C ABI: R-MAT-24 and Erdos-Renyi-24 (CSR build, pull, push, fixed point),
batched personalized PageRank at R-MAT-22 with k = 64, the 256-graph uneven
batch, and the learned warm start (200 R-MAT-18 training graphs, an R-MAT-24
evaluation graph).

Bars (the same as at small sizes, in test_gpu_parity and the others):
* CSR arrays bit-exact against the oracle's restatement of the reference's
  radix sort (itself pinned by the reference's digests and numpy's stable
  sorts at small sizes), for -o 0 and -o 1;
* three iterations at epsilon 0 (every iteration does full work) on the same
  CSR: the default (binned) pull and the gather pull with every row exact
  bit-identical to the oracle's sequential pull, the gather pull's default
  split within 1e-5 relative, fixed-point pull and push
  bit-identical to the oracle and to each other, fp32 push within 1e-5;
* batched PPR: every column converged, its iteration count within one of
  its single-source run's and its L1 distance to that run at most 1e-6;
  four columns within 1e-5 L1 of FP64 PPR;
* uneven batch: sampled graphs bit-identical to their own single-graph runs
  with the same row split (chunk trees above the hub threshold, no chains);
* learned start: falls within 1e-5 L1 of the FP64 fixed point after reaching
  an L1 change < 1e-6, like the uniform start.
The oracle runs on the host's cores (at most 16 threads).
"""
import os

import numpy as np
import pytest

import opengraph_amd as og
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
ALL_EXACT = 0xFFFFFFFF


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


def _device_graph(kind, scale, ef, seed, sort, a=0.57, b=0.19, c=0.19):
    E = ef << scale
    if kind == "rmat":
        ds, dd = og.rmat_device(scale, E, a, b, c, seed, True)
    else:
        ds, dd = og.er_device(1 << scale, E, seed)
    g = og.Graph.build(ds, dd, E, sort)
    src, dst = ds.to_numpy(np.uint32, E), dd.to_numpy(np.uint32, E)
    ds.free()
    dd.free()
    return g, src, dst


def _csr_equal(g, o, label):
    for k in po.CSR_FIELDS:
        got = g.array(k)
        assert np.array_equal(got, o[k]), (label, k)


@pytest.fixture(scope="module")
def rmat24():
    """R-MAT scale 24, edge factor 16, a/b/c 0.57/0.19/0.19, seed 1 (the
    benchmark's graph), sort -o 0, and the oracle's CSR of the same edges."""
    g, s, d = _device_graph("rmat", 24, 16, 1, 0)
    print(f"R-MAT-24: N={g.num_vertices} E={g.num_edges}", flush=True)
    o = po.build_csr(s, d, 0, THREADS)
    yield g, o, s, d
    g.free()


def _run(g, code, iters, **kw):
    pr = og.PageRank(g, variant=code, max_iter=iters, epsilon=0.0, **kw)
    res = pr.run()
    x = pr.ranks()
    pr.free()
    return res, x


def test_c2_rmat24_csr_bitexact(rmat24):
    g, o, s, d = rmat24
    assert g.num_vertices == o.num_vertices and g.num_edges == o.num_edges == 16 << 24
    _csr_equal(g, o, "rmat24 -o 0")


def test_c2_rmat24_csr_bitexact_sort1(rmat24):
    """-o 1: by destination, then by source (radixsort.c:282-335)."""
    _, _, s, d = rmat24
    E = len(s)
    ds, dd = og.DeviceArray.from_numpy(s), og.DeviceArray.from_numpy(d)
    g1 = og.Graph.build(ds, dd, E, 1)
    ds.free()
    dd.free()
    o1 = po.build_csr(s, d, 1, THREADS)
    _csr_equal(g1, o1, "rmat24 -o 1")
    g1.free()


def test_c2_rmat24_pull_push_fixed_point(rmat24):
    g, o, _, _ = rmat24
    it = 3
    ref0 = po.pagerank(o, 0, 0.0, it, THREADS)
    ref2 = po.pagerank(o, 2, 0.0, it, THREADS)
    assert ref0["iterations"] == it
    # the default (binned) pull and the gather pull with every row in the
    # reference's order: the same bits
    for kw in ({}, {"hub_threshold": ALL_EXACT, "binned": 2}):
        res, x = _run(g, 0, it, **kw)
        assert res.iterations == it
        assert np.array_equal(x.view(np.uint32), ref0["ranks"].view(np.uint32)), kw
    # the gather pull's default split (chunk trees below 4096 in-edges)
    res, x = _run(g, 0, it, binned=2)
    rel = float(np.max(np.abs(x - ref0["ranks"]) / ref0["ranks"]))
    print("gather pull default split max rel", rel, flush=True)
    assert rel <= 1e-5
    # fixed point: integer sums, bit-exact, pull == push
    _, x2 = _run(g, 2, it)
    assert np.array_equal(x2.view(np.uint32), ref2["ranks"].view(np.uint32))
    _, x3 = _run(g, 3, it)
    assert np.array_equal(x3.view(np.uint32), ref2["ranks"].view(np.uint32))
    # fp32 push: atomic adds in no fixed order (as the reference's omp atomic)
    _, x1 = _run(g, 1, it)
    rel1 = float(np.max(np.abs(x1 - ref0["ranks"]) / ref0["ranks"]))
    ref1 = po.pagerank(o, 1, 0.0, it, THREADS)["ranks"]
    rel_ref = float(np.max(np.abs(ref1 - ref0["ranks"]) / ref0["ranks"]))
    print("push max rel vs sequential pull", rel1, "(the oracle's own push:", rel_ref, ")", flush=True)
    assert rel1 <= 1e-5


@pytest.fixture(scope="module")
def er24():
    g, s, d = _device_graph("er", 24, 16, 2, 0)
    print(f"ER-24: N={g.num_vertices} E={g.num_edges}", flush=True)
    o = po.build_csr(s, d, 0, THREADS)
    yield g, o
    g.free()


def test_c2_er24_csr_and_pagerank(er24):
    g, o = er24
    _csr_equal(g, o, "er24 -o 0")
    it = 3
    ref0 = po.pagerank(o, 0, 0.0, it, THREADS)
    ref2 = po.pagerank(o, 2, 0.0, it, THREADS)
    _, x = _run(g, 0, it, binned=2)
    rel = float(np.max(np.abs(x - ref0["ranks"]) / ref0["ranks"]))
    print("ER-24 gather pull max rel", rel, flush=True)
    assert rel <= 1e-5
    for kw in ({}, {"hub_threshold": ALL_EXACT, "binned": 2}):
        _, x = _run(g, 0, it, **kw)
        assert np.array_equal(x.view(np.uint32), ref0["ranks"].view(np.uint32)), kw
    for code in (2, 3):
        _, x = _run(g, code, it)
        assert np.array_equal(x.view(np.uint32), ref2["ranks"].view(np.uint32)), code
    _, x1 = _run(g, 1, it)
    assert float(np.max(np.abs(x1 - ref0["ranks"]) / ref0["ranks"])) <= 1e-5


def test_c2_rmat24_converged_vs_fp64(rmat24):
    """To the stop rule (tol 1e-6, uniform start, L1 stop): within 1e-5 L1 of
    the FP64 fixed point."""
    g, o, _, _ = rmat24
    pr = og.PageRank(g, 0, 200, 1e-6, stop_rule=1, init_mode=1)
    res = pr.run()
    x = pr.ranks()
    pr.free()
    x64, it64, l1 = po.pagerank_fp64(o, 1.0, 1e-11, 400, THREADS)
    err = float(np.abs(x.astype(np.float64) - x64).sum())
    print("R-MAT-24 uniform start: iterations", res.rows, "L1 vs fp64", err, "(fp64", it64, l1, ")", flush=True)
    assert res.error_total < 1e-6 and err <= 1e-5


def test_c3_batched_ppr_rmat22_k64():
    """R-MAT scale 22 (seed 3), k = 64 sources uniform over the vertices
    (seed 4), tol 1e-6 per column: every column converges, within one
    iteration of its single-source run (default rows: chunk trees above 16
    in-edges, as the batch's 512-edge pieces are) and within 2 * tol *
    d / (1 - d) L1 of it (each run stops once an iteration moves the ranks
    by less than tol, which leaves it within tol * d / (1 - d) of the fixed
    point: the update contracts by d = 0.85 in L1), and four columns within
    1e-5 L1 of FP64 PPR.  The single-source run with every
    row in the reference's order is reported beside them: its long fp32
    running sums drift further from FP64 at this scale."""
    g, s, d = _device_graph("rmat", 22, 16, 3, 0)
    src = np.random.default_rng(4).integers(0, g.num_vertices, size=64, dtype=np.uint32)
    p = og.Ppr(g, src, 100, 1e-6)
    r = p.run()
    X = p.ranks()
    iters, _ = p.columns()
    p.free()
    assert r.converged == 64
    worst = 0.0
    for j in range(64):
        pr = og.PageRank(g, 0, 100, 1e-6, ppr_source=int(src[j]))
        r1 = pr.run()
        x = pr.ranks()
        pr.free()
        l1 = float(np.abs(X[:, j].astype(np.float64) - x).sum())
        worst = max(worst, l1)
        assert abs(int(iters[j]) - int(r1.rows)) <= 1, j
        assert l1 <= 2 * 1e-6 * 0.85 / 0.15, (j, l1)
    print("batched vs single-source worst L1", worst, flush=True)
    o = po.build_csr(s, d, 0, THREADS)
    for j in range(4):
        x64, it64, _ = po.ppr_fp64(o, int(src[j]), 1e-10, 500)
        l1 = float(np.abs(X[:, j].astype(np.float64) - x64).sum())
        pr = og.PageRank(g, 0, 100, 1e-6, ppr_source=int(src[j]), hub_threshold=ALL_EXACT)
        pr.run()
        xe = pr.ranks()
        pr.free()
        l1e = float(np.abs(xe.astype(np.float64) - x64).sum())
        print("column", j, "L1 vs fp64: batched", l1, "single-source in the reference's order", l1e,
              "fp64 iterations", it64, flush=True)
        assert l1 <= 1e-5
    g.free()


def test_c4_uneven_batch_256():
    """256 R-MAT graphs, scale U{12..20}, edge factor U{8..32}, a U[0.45,
    0.65], b = c = (1 - a - 0.05)/2, master seed 5 (the benchmark's batch):
    one hipGraph launch with autotuned thresholds; every 16th graph, and the
    largest, bit-identical to its own single-graph run with that threshold and
    no chains (exact_above 0: the batch sums every long row in chunks)."""
    rng = np.random.default_rng(5)
    gs = []
    for i in range(256):
        sc = int(rng.integers(12, 21)); ef = int(rng.integers(8, 33)); a = float(rng.uniform(0.45, 0.65))
        b = (1 - a - 0.05) / 2
        E = ef << sc
        ds, dd = og.rmat_device(sc, E, a, b, b, 5000 + i, True)
        gs.append(og.Graph.build(ds, dd, E, 0))
        ds.free()
        dd.free()
    m = og.Multi(gs, 100, 1e-6)
    m.run(2)
    t = m.run(2)                              # a replay
    largest = int(np.argmax([g.num_edges for g in gs]))
    picks = sorted(set(list(range(0, 256, 16)) + [largest]))
    for i in picks:
        res, x = m.result(i)
        pr = og.PageRank(gs[i], 0, 100, 1e-6, hub_threshold=m.hub_threshold(i), exact_above=0, binned=2)
        r1 = pr.run()
        x1 = pr.ranks()
        pr.free()
        assert res.iterations == r1.iterations and res.rows == r1.rows, i
        assert np.array_equal(x.view(np.uint32), x1.view(np.uint32)), i
    print("256 graphs in one hipGraph:", t, "s; thresholds", sorted({m.hub_threshold(i) for i in range(256)}),
          flush=True)
    m.free()
    for g in gs:
        g.free()


def test_c5_learned_start_rmat24():
    """Training set: 200 R-MAT-18 graphs (a U[0.45, 0.65], edge factor
    U{8..32}, seed 6), converged ranks from this PageRank; evaluation graph:
    R-MAT-24 seed 103 (the smallest of the benchmark's held-out graphs).
    The learned start is used (not rejected), the run reaches an L1 change
    < 1e-6, and its ranks are within 1e-5 L1 of the FP64 fixed point."""
    rng = np.random.default_rng(6)
    m = og.Mlp(6)
    for i in range(200):
        a = float(rng.uniform(0.45, 0.65)); ef = int(rng.integers(8, 33)); b = (1 - a - 0.05) / 2
        E = ef << 18
        ds, dd = og.rmat_device(18, E, a, b, b, 6000 + i, True)
        g = og.Graph.build(ds, dd, E, 0)
        ds.free()
        dd.free()
        assert m.add_graph(g, 1e-9, 2000) < 1e-9
        g.free()
    loss0, _ = m.train(1, 65536, 0.0)
    loss, _ = m.train(8, 65536, 3e-3, 1)
    print("training loss", loss0, "->", loss, flush=True)
    assert loss < loss0
    r = np.random.default_rng(103)
    a = float(r.uniform(0.45, 0.65)); ef = int(r.integers(8, 33)); b = (1 - a - 0.05) / 2
    g, s, d = _device_graph("rmat", 24, ef, 103, 0, a, b, b)
    res = {}
    for name in ("uniform", "learned"):
        pr = og.PageRank(g, 0, 1000, 1e-6, stop_rule=1, init_mode=1)
        if name == "learned":
            pr.set_init_learned(m)
            assert not pr.init_fallback()
        rr = pr.run()
        res[name] = (rr, pr.ranks())
        pr.free()
    o = og_csr_from_graph(g)
    x64, it64, _ = po.pagerank_fp64(o, 1.0, 1e-11, 500, THREADS)
    for name, (rr, x) in res.items():
        err = float(np.abs(x.astype(np.float64) - x64).sum())
        print(name, "iterations", rr.rows, "last L1 change", rr.error_total, "L1 vs fp64", err, flush=True)
        assert rr.error_total < 1e-6 and err <= 1e-5, name
    m.free()
    g.free()


def og_csr_from_graph(g):
    """The oracle's CSR object over a device graph's arrays (the GPU build is
    pinned bit-exact against the oracle's in the tests above)."""
    arrays = g.arrays()
    return po.CSR(arrays, g.num_vertices, g.num_edges)
