"""Learned warm start on the GPU, through the C ABI.

Bars: features within 1e-5 relative of the oracle's double-precision ones
(fp32 log1p and a sum in another order); fp32 inference within 1e-5 of the
oracle's double forward; the MFMA inference within 1e-4 of a numpy model of
the same bf16 roundings (and within bf16 tolerance of fp32); the hand-written
backward within 1e-3 of the oracle gradient (itself pinned by finite
differences on the CPU); training lowers the loss; the warm-started run
stops at the same tolerance on the same fixed point (FP64 oracle L1 <= 1e-5),
the trained start is closer to it than the uniform and untrained ones, and
(with the class correction) takes at most 0.8 times the uniform start's
iterations and no more than the untrained network's start; the corrected
start balances every class's mass under one iteration (fp64 check from
the edge list, 1e-5 relative).
"""
import numpy as np
import pytest

import opengraph_amd as og
from conftest import golden_input
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


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


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(og.MLP_PARAMS, np.float32)
    p[0:256] = rng.normal(0, np.sqrt(2 / 4), 256)
    p[256:320] = rng.normal(0, 0.1, 64)
    p[320:4416] = rng.normal(0, np.sqrt(2 / 64), 4096)
    p[4416:4480] = rng.normal(0, 0.1, 64)
    p[4480:4544] = rng.normal(0, np.sqrt(1 / 64), 64)
    p[4544] = 0.2
    return p


def _bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def test_features_match_oracle(rmat16):
    g, o = rmat16
    pr = og.PageRank(g, 0, 10, 1e-6)
    f = pr.features()
    pr.free()
    ref = po.mlp_features(o)
    np.testing.assert_allclose(f[:, :2], ref[:, :2], rtol=1e-6)
    np.testing.assert_allclose(f[:, 2], ref[:, 2], rtol=1e-5, atol=1e-6)
    assert np.all(f[:, 3] == ref[:, 3])


def test_forward_fp32_and_mfma(rmat16):
    _, o = rmat16
    feat = po.mlp_features(o)[:20000]
    p = _rand_params(3)
    m = og.Mlp(1)
    m.set_params(p)
    ref = po.mlp_forward(p, feat)
    out = m.forward(feat, mfma=False)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6)
    # MFMA path: bf16 h1 and W2, fp32 accumulation -- numpy model of exactly that
    W1, b1 = p[0:256].reshape(64, 4), p[256:320]
    W2, b2 = p[320:4416].reshape(64, 64), p[4416:4480]
    W3, b3 = p[4480:4544], p[4544]
    h1 = np.maximum(feat.astype(np.float64) @ W1.T.astype(np.float64) + b1, 0).astype(np.float32)
    z2 = _bf16(h1).astype(np.float64) @ _bf16(W2).T.astype(np.float64) + b2
    z3 = np.maximum(z2, 0) @ W3 + b3
    model = np.where(z3 > 20, z3, np.log1p(np.exp(np.minimum(z3, 20))))
    mf = m.forward(feat, mfma=True)
    m.free()
    np.testing.assert_allclose(mf, model, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(mf, ref, rtol=5e-2, atol=5e-3)


def test_backward_matches_oracle(rmat16):
    _, o = rmat16
    rng = np.random.default_rng(5)
    feat = po.mlp_features(o)[rng.choice(o.num_vertices, 6000, replace=False)]
    p = _rand_params(4)
    y = (po.mlp_forward(p, feat) * rng.uniform(0.5, 1.5, len(feat))).astype(np.float32)
    m = og.Mlp(1)
    m.set_params(p)
    loss, g = m.grad(feat, y)
    m.free()
    lref, gref = po.mlp_grad(p, feat, y)
    assert loss == pytest.approx(lref, rel=1e-5)
    scale = np.abs(gref).max()
    np.testing.assert_allclose(g, gref, rtol=1e-3, atol=1e-4 * scale)


def _graphs(scale, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        a = float(rng.uniform(0.45, 0.65))
        b = (1 - a - 0.05) / 2
        ef = int(rng.integers(8, 33))
        s, d = og.rmat_host(scale, ef << scale, a, b, b, seed * 1000 + i, True)
        out.append((s, d, og.Graph.from_edges(s, d, 0)))
    return out


def test_training_and_warm_start():
    train = _graphs(12, 6, 6)
    m = og.Mlp(7)
    untrained = og.Mlp(7)
    for _, _, g in train:
        l1 = m.add_graph(g, 1e-10, 1000)
        assert l1 < 1e-10
    assert m.data_size() == sum(g.num_vertices for _, _, g in train)
    loss0, _ = m.train(1, 16384, 0.0)
    loss, sec = m.train(30, 16384, 3e-3)
    print("loss", loss0, "->", loss, "seconds", sec, flush=True)
    assert loss < 0.7 * loss0
    s, d, g = _graphs(14, 1, 101)[0]
    o = po.build_csr(s, d, 0, 8)
    x64, _, _ = po.pagerank_fp64(o, 1.0, 1e-14, 3000, 8)
    N = g.num_vertices
    res = {}
    for name, net, wc in (("uniform", None, 0), ("untrained", untrained, 0), ("trained", m, 0),
                          ("trained, no class correction", m, 1)):
        pr = og.PageRank(g, 0, 500, 1e-6, stop_rule=1, init_mode=1, warm_classes=wc)
        if net is not None:
            pr.set_init_learned(net)
        r = pr.run()
        x = pr.ranks()
        pr.free()
        res[name] = r.rows
        err = np.abs(x.astype(np.float64) - x64).sum()
        print(name, "iterations", r.rows, "L1 vs fp64", err, flush=True)
        assert err <= 1e-5, name
    # start-vector error against the fixed point
    starts = {}
    for name, net in (("untrained", untrained), ("trained", m)):
        pr = og.PageRank(g, 0, 1, 1e-6, stop_rule=1, init_mode=1)
        pr.set_init_learned(net)
        f = pr.features()
        pr.free()
        q = net.forward(f, mfma=False).astype(np.float64)
        dang = o["f_out_degree"] == 0
        K = o["i_out_degree"] == 0                    # no in-edges: fixed point (1-d)/N
        qU = np.where(K, 0.0, q)
        shape = qU / qU.sum()
        SU = 0.15 * (1 - 0.15 * K.mean() - 0.85 * (K & dang).mean()) / (0.15 + 0.85 * shape[dang].sum())
        x0 = np.where(K, 0.15 / N, shape * SU)
        starts[name] = np.abs(x0 - x64).sum()
    starts["uniform"] = np.abs(np.full(N, 1.0 / N) - x64).sum()
    print("start L1 to the fixed point", starts, "iterations", res, flush=True)
    assert starts["trained"] < starts["untrained"] and starts["trained"] < starts["uniform"]
    # what the trained network gains: a start closer to the fixed point than
    # the untrained network's (above), fewer iterations than the uniform
    # start, and no more than the same class-corrected start from the
    # untrained network (the correction removes most of the coarse error of
    # either start, so the iteration counts end close)
    assert res["trained"] <= 0.8 * res["uniform"], res
    assert res["trained"] <= res["untrained"], res
    m.free(); untrained.free()
    for _, _, g in train:
        g.free()
    g.free()


def _classes(indeg, outdeg, nin):
    """Class 2 i + o: i the in-degree class among nin (0 without in-edges,
    else 1 + min(floor(log2 in-degree), nin - 2)), o = 1 with out-edges."""
    l = np.floor(np.log2(np.maximum(indeg, 1))).astype(np.int64)
    ci = np.where(indeg == 0, 0, 1 + np.minimum(l, nin - 2))
    return 2 * ci + (outdeg > 0)


@pytest.mark.parametrize("nin", [16, 4])
def test_class_correction_balances_class_masses(nin):
    """The corrected start's defining property: the mass of every class
    equals the class's mass after one application of the iteration,
    x0 -> 0.15 + 0.85 A (x0 / out-degree) (computed here in fp64 from the
    edge list); the uncorrected start does not have it."""
    s, d, g = _graphs(14, 1, 77)[0]
    N = g.num_vertices
    indeg = np.bincount(d, minlength=N)
    outdeg = np.bincount(s, minlength=N)
    c = _classes(indeg, outdeg, nin)
    net = og.Mlp(7)

    def balance(wc):
        pr = og.PageRank(g, 0, 1, 1e-6, stop_rule=1, init_mode=1, warm_classes=wc)
        pr.set_init_learned(net)
        assert not pr.init_fallback()
        x0 = pr.get_init().astype(np.float64)
        pr.free()
        contrib = np.where(outdeg > 0, x0 / np.maximum(outdeg, 1), 0.0)
        fx = 0.15 + 0.85 * np.bincount(d, weights=contrib[s], minlength=N)
        m0 = np.bincount(c, weights=x0, minlength=2 * nin)
        m1 = np.bincount(c, weights=fx, minlength=2 * nin)
        keep = m0 > 0
        return np.abs(m1[keep] - m0[keep]) / m0[keep]

    rel = balance(nin)
    print("classes", int((np.bincount(c) > 0).sum()), "max relative class-mass change", rel.max(), flush=True)
    assert rel.max() < 1e-5
    assert balance(1).max() > 1e-3          # the prediction alone is not balanced
    net.free()
    g.free()


def test_rejected_start_falls_back_to_uniform():
    """A learned start with NaN parameters, or whose prediction has no mass
    (softplus underflow), and a caller's start with a NaN or a negative entry
    are rejected: the run equals the uniform-start run bit for bit."""
    s, d, g = _graphs(12, 1, 55)[0]
    N = g.num_vertices
    pr = og.PageRank(g, 0, 200, 1e-6, stop_rule=1, init_mode=1)
    ru = pr.run()
    xu = pr.ranks()
    assert not pr.init_fallback()
    pr.free()
    nan_net = og.Mlp(3)
    nan_net.set_params(np.full(nan_net.params().size, np.nan, np.float32))
    zero_net = og.Mlp(3)
    p = np.zeros(zero_net.params().size, np.float32)
    p[-1] = -200.0                     # b3: softplus(-200) == 0 in fp32 on every vertex
    zero_net.set_params(p)
    good_net = og.Mlp(3)
    bad_nan = np.full(N, 1.0, np.float32)
    bad_nan[N // 2] = np.nan
    bad_neg = np.full(N, 1.0, np.float32)
    bad_neg[1] = -0.5
    cases = [("nan params", nan_net, None), ("zero mass", zero_net, None),
             ("nan entry", None, bad_nan), ("negative entry", None, bad_neg),
             ("all zero", None, np.zeros(N, np.float32))]
    for name, net, vec in cases:
        pr = og.PageRank(g, 0, 200, 1e-6, stop_rule=1, init_mode=1)
        if net is not None:
            pr.set_init_learned(net)
        else:
            pr.set_init(vec)
        assert pr.init_fallback(), name
        r = pr.run()
        x = pr.ranks()
        pr.free()
        print(name, "iterations", r.rows, "uniform", ru.rows, flush=True)
        assert r.rows == ru.rows and r.iterations == ru.iterations, name
        assert np.array_equal(x.view(np.uint32), xu.view(np.uint32)), name
    # an accepted start is used (and differs from the uniform run)
    pr = og.PageRank(g, 0, 200, 1e-6, stop_rule=1, init_mode=1)
    pr.set_init_learned(good_net)
    assert not pr.init_fallback()
    pr.run()
    assert not np.array_equal(pr.ranks().view(np.uint32), xu.view(np.uint32))
    pr.free()
    for net in (nan_net, zero_net, good_net):
        net.free()
    g.free()
