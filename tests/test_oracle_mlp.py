"""Pins of the learned warm start's CPU oracle (og_oracle_mlp.c).

The features are checked against numpy restatements on golden graphs; the
double-precision gradient against central finite differences of the loss
(with a planted wrong sign caught); the forward pass against a numpy
restatement of the network.
"""
import numpy as np
import pytest

from conftest import golden_input
from oracle import pyoracle as po

P = po.MLP_PARAMS


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    p[0:256] = rng.normal(0, np.sqrt(2 / 4), 256)
    p[256:320] = rng.normal(0, 0.1, 64)
    p[320:4416] = rng.normal(0, np.sqrt(2 / 64), 4096)
    p[4416:4480] = rng.normal(0, 0.1, 64)
    p[4480:4544] = rng.normal(0, np.sqrt(1 / 64), 64)
    p[4544] = 0.3
    return p


def _np_forward(p, f):
    p = p.astype(np.float64)
    W1, b1 = p[0:256].reshape(64, 4), p[256:320]
    W2, b2 = p[320:4416].reshape(64, 64), p[4416:4480]
    W3, b3 = p[4480:4544], p[4544]
    h1 = np.maximum(f.astype(np.float64) @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    z = h2 @ W3 + b3
    return np.where(z > 30, z, np.log1p(np.exp(np.minimum(z, 30))))


@pytest.mark.parametrize("name", ["graphbrew", "enron"])
def test_features_closed_form(golden, name):
    s, d = golden_input(name, golden)
    g = po.build_csr(s, d, 0, 4)
    f = po.mlp_features(g)
    N = g.num_vertices
    ind = np.bincount(d, minlength=N)
    outd = np.bincount(s, minlength=N)
    li = np.log1p(ind.astype(np.float64))
    mean = np.bincount(d, weights=li[s], minlength=N) / np.maximum(ind, 1)
    np.testing.assert_allclose(f[:, 0], np.log1p(ind), rtol=1e-6)
    np.testing.assert_allclose(f[:, 1], np.log1p(outd), rtol=1e-6)
    np.testing.assert_allclose(f[:, 2], mean, rtol=1e-6, atol=1e-7)
    assert np.all(f[:, 3] == np.float32(1.0 / N))


def test_forward_matches_numpy():
    rng = np.random.default_rng(3)
    f = rng.uniform(0, 5, (500, 4)).astype(np.float32)
    p = _params(1)
    np.testing.assert_allclose(po.mlp_forward(p, f), _np_forward(p, f), rtol=1e-12)


def test_gradient_finite_differences():
    rng = np.random.default_rng(4)
    f = rng.uniform(0, 4, (100, 4)).astype(np.float32)
    p = _params(2)
    y = (po.mlp_forward(p, f) * rng.uniform(0.5, 1.5, 100)).astype(np.float32)
    loss, g = po.mlp_grad(p, f, y)
    assert loss == pytest.approx(po.mlp_loss(p, f, y), rel=1e-12)
    idx = np.concatenate([rng.choice(256, 8), 256 + rng.choice(64, 6), 320 + rng.choice(4096, 16),
                          4416 + rng.choice(64, 6), 4480 + rng.choice(64, 6), [4544]])
    wrong = 0
    for i in idx:
        h = 1e-5 * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        hh = (float(pp[i]) - float(pm[i])) / 2     # the step fp32 actually took
        fd = (po.mlp_loss(pp, f, y) - po.mlp_loss(pm, f, y)) / (2 * hh)
        assert fd == pytest.approx(g[i], rel=2e-2, abs=2e-6), i
        wrong += abs(fd + g[i]) > 1e-3 * max(abs(fd), 1e-6)
    assert wrong > 0        # a planted sign flip of the gradient would be caught
