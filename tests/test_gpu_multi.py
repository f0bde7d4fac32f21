"""Many graphs in one launch and the autotuner, through the C ABI.

Bar: every graph's result in the batched modes (all graphs per launch, and
the same captured in one hipGraph) is bit-identical to its own single-graph
run with the same row-class threshold -- the kernels and plans are the same
-- in ranks, iteration count and active counts (the fp64 error sum to 1e-12
relative: its partials are folded in another order); the single-graph
runs are themselves pinned against the oracle in test_gpu_parity.  The
hipGraph is replayed several times with eager work between the replays.
"""
import numpy as np
import pytest

import opengraph_amd as og
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    og.init(0)


def _params(n, seed):
    """Uneven batch: scale, edge factor, a (b = c = (1 - a - 0.05) / 2)."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(6, 14)), int(rng.integers(8, 33)), float(rng.uniform(0.45, 0.65)))
            for _ in range(n)]


@pytest.fixture(scope="module")
def batch():
    gs, edges = [], []
    for i, (sc, ef, a) in enumerate(_params(12, 5)):
        b = (1.0 - a - 0.05) / 2
        s, d = og.rmat_host(sc, ef << sc, a, b, b, 1000 + i, True)
        gs.append(og.Graph.from_edges(s, d, 0))
        edges.append((s, d))
    # a tiny graph with a vertex of no in- or out-edges
    s, d = np.array([0, 1, 1], np.uint32), np.array([1, 0, 3], np.uint32)
    gs.append(og.Graph.from_edges(s, d, 0))
    edges.append((s, d))
    yield gs, edges
    for g in gs:
        g.free()


def _single(g, hub, iters, eps):
    pr = og.PageRank(g, 0, iters, eps, hub_threshold=hub, binned=2)
    res = pr.run()
    x = pr.ranks()
    act, err = pr.rows()
    pr.free()
    return res, x, act, err


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_multi_matches_single_graph_runs(batch, mode):
    gs, _ = batch
    hubs = [8, 16, 32, 64] * 4
    m = og.Multi(gs, 60, 1e-6, hub_thresholds=hubs[:len(gs)])
    for rep in range(2 if mode == 2 else 1):
        t = m.run(mode)
        print("mode", mode, "replay", rep, "seconds", t, flush=True)
        for i, g in enumerate(gs):
            res, x = m.result(i)
            r1, x1, act1, err1 = _single(g, hubs[i], 60, 1e-6)
            assert res.iterations == r1.iterations and res.rows == r1.rows, i
            assert np.array_equal(x.view(np.uint32), x1.view(np.uint32)), i
            assert res.active == r1.active, i
            # the fp64 error sum folds the same per-vertex terms in another tree
            assert res.error_total == pytest.approx(r1.error_total, rel=1e-12, abs=1e-300), i
        # eager work between replays
        og.synchronize()
    m.free()


def test_multi_epsilon_zero_counts_settled_rows(batch):
    """epsilon 0: every vertex is active in every iteration (|delta| >= 0),
    including the rows without in-edges that the batched pull leaves out
    from the third iteration on; ranks as the single-graph run's."""
    gs, _ = batch
    m = og.Multi(gs, 6, 0.0, hub_thresholds=[16] * len(gs))
    m.run(1)
    for i, g in enumerate(gs):
        res, x = m.result(i)
        r1, x1, act1, _ = _single(g, 16, 6, 0.0)
        assert res.iterations == r1.iterations and res.active == g.num_vertices == r1.active, i
        assert np.array_equal(x.view(np.uint32), x1.view(np.uint32)), i
    m.free()


def test_multi_results_match_oracle(batch):
    gs, edges = batch
    m = og.Multi(gs, 40, 1e-6, hub_threshold=0xFFFFFFFF)      # every row in CSR order
    m.run(2)
    for i, (s, d) in enumerate(edges[:6]):
        o = po.build_csr(s, d, 0, 4)
        ref = po.pagerank(o, 0, 1e-6, 40, 4)
        res, x = m.result(i)
        assert np.array_equal(x.view(np.uint32), ref["ranks"].view(np.uint32)), i
        assert res.iterations == ref["iterations"], i
    m.free()


def test_autotuned_thresholds_are_candidates(batch):
    gs, _ = batch
    m = og.Multi(gs[:5], 20, 1e-6)
    chosen = [m.hub_threshold(i) for i in range(5)]
    assert all(c in (8, 16, 32, 64, 128) for c in chosen), chosen
    m.run(1)
    for i in range(5):
        res, x = m.result(i)
        r1, x1, _, _ = _single(gs[i], chosen[i], 20, 1e-6)
        assert np.array_equal(x.view(np.uint32), x1.view(np.uint32))
    m.free()


@pytest.fixture(scope="module")
def singles(batch):
    """Each graph's single-graph run (threshold 16, 30 iterations)."""
    gs, _ = batch
    return [_single(g, 16, 30, 1e-6) for g in gs]


def _check_against_singles(m, gs, singles):
    for i in range(len(gs)):
        res, x = m.result(i)
        r1, x1, act1, err1 = singles[i]
        assert res.iterations == r1.iterations and res.active == r1.active, i
        assert np.array_equal(x.view(np.uint32), x1.view(np.uint32)), i
        assert res.error_total == pytest.approx(r1.error_total, rel=1e-12, abs=1e-300), i


def test_every_tile_shape_matches_single_graph_runs(batch, singles):
    """Block size, chunks per wavefront and row batch change neither a row's
    sum nor a chunk's: every compiled shape gives each graph its
    single-graph ranks, bit for bit."""
    gs, _ = batch
    shapes = og.Multi.tile_shapes()
    assert len(shapes) == 18 and (256, 4, 8) in shapes
    for shape in shapes:
        m = og.Multi(gs, 30, 1e-6, hub_thresholds=[16] * len(gs), tiles=[shape] * len(gs))
        assert all(m.tile(i) == shape for i in range(len(gs)))
        m.run(1)
        print("tile", shape, flush=True)
        _check_against_singles(m, gs, singles)
        m.free()


def test_mixed_tile_shapes_in_one_hipgraph(batch, singles):
    """Graphs of different shapes: one pull launch per shape in each
    iteration, captured in one hipGraph and replayed with eager work
    between the replays."""
    gs, _ = batch
    shapes = og.Multi.tile_shapes()
    tiles = [shapes[(5 * i) % len(shapes)] for i in range(len(gs))]
    m = og.Multi(gs, 30, 1e-6, hub_thresholds=[16] * len(gs), tiles=tiles)
    for rep in range(3):
        m.run(2)
        _check_against_singles(m, gs, singles)
        og.synchronize()
    m.free()


def test_autotuned_tile_shapes(batch, singles):
    """Tuned per size class (graphs of the same ceil(log2 N)) among the
    compiled shapes; given shapes are kept; results as single runs."""
    gs, _ = batch
    shapes = og.Multi.tile_shapes()
    tiles = [(0, 0, 0)] * len(gs)
    tiles[0] = (128, 2, 0)
    m = og.Multi(gs, 30, 1e-6, hub_thresholds=[16] * len(gs), tiles=tiles)
    got = [m.tile(i) for i in range(len(gs))]
    print("tuned", got, flush=True)
    assert got[0] == (128, 2, 0)
    assert all(t in shapes for t in got)
    cls = [int(np.ceil(np.log2(max(g.num_vertices, 1)))) for g in gs]
    for i in range(1, len(gs)):
        for j in range(1, len(gs)):
            if cls[i] == cls[j]:
                assert got[i] == got[j], (i, j)
    m.run(2)
    _check_against_singles(m, gs, singles)
    m.free()
    with pytest.raises(RuntimeError):
        og.Multi(gs[:2], 10, 1e-6, hub_thresholds=[16, 16], tiles=[(192, 4, 8), (256, 4, 8)])


def test_tuner_reports_every_candidate(batch):
    gs, _ = batch
    best, ms = og.tune_hub_threshold(gs[0], (4, 32, 256), 3)
    assert best in (4, 32, 256)
    assert np.all(ms > 0) and ms[[4, 32, 256].index(best)] == ms.min()


def test_multi_rejects_other_variants(batch):
    gs, _ = batch
    o = og.PrOptions()
    og.load().ogpu_pr_default_options(og.C.byref(o))
    o.variant = 2
    h = og.C.c_void_p()
    hs = (og.C.c_void_p * 1)(gs[0].h.value if isinstance(gs[0].h, og.C.c_void_p) else gs[0].h)
    assert og.load().ogpu_multi_create(hs, 1, og.C.byref(o), None, og.C.byref(h)) != 0
