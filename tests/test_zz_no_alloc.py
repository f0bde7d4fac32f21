"""ogpu_mlp_grad_noalloc against ogpu_mlp_grad, through the C ABI.

The no-allocation entry point runs the same gradient kernels with their
block partials in the library's static device array instead of a fresh
allocation, so its gradient and loss must equal ogpu_mlp_grad's bit for bit
(the partial layout and the fixed-order reduction are the same).  Two calls
in a row, the second over fewer rows, so the second reads only partials the
second call wrote; then the size limit is an error, not a launch.
"""
import ctypes as C

import numpy as np
import pytest

import opengraph_amd as og

pytestmark = pytest.mark.gpu


def _rows(n, seed):
    rng = np.random.default_rng(seed)
    feat = np.empty((n, 4), np.float32)
    feat[:, 0] = np.log1p(rng.integers(0, 2000, n)).astype(np.float32)
    feat[:, 1] = np.log1p(rng.integers(0, 2000, n)).astype(np.float32)
    feat[:, 2] = rng.uniform(0.0, 6.0, n).astype(np.float32)
    feat[:, 3] = np.float32(1.0 / (1 << 18))
    y = rng.gamma(1.0, 1.0, n).astype(np.float32)
    return feat, y


def _noalloc(lib, m, feat, y):
    n = len(y)
    d_feat = og.DeviceArray.from_numpy(np.ascontiguousarray(feat))
    d_y = og.DeviceArray.from_numpy(y)
    d_grad = og.DeviceArray(og.MLP_PARAMS * 4)
    d_loss = og.DeviceArray(8)
    og.check(lib.ogpu_mlp_grad_noalloc(m.h, d_feat.ptr, d_y.ptr, n, d_grad.ptr, d_loss.ptr))
    og.check(lib.ogpu_synchronize())
    grad = d_grad.to_numpy(np.float32, og.MLP_PARAMS)
    loss_sum = d_loss.to_numpy(np.float64, 1)[0]
    for d in (d_feat, d_y, d_grad, d_loss):
        d.free()
    return grad, loss_sum


def test_noalloc_gradient_equals_allocating_gradient():
    og.init(0)
    lib = og.load()
    m = og.Mlp(11)
    limit = 16384
    for n, seed in ((limit, 1), (3001, 2)):
        feat, y = _rows(n, seed)
        want_loss, want_grad = m.grad(feat, y)
        grad, loss_sum = _noalloc(lib, m, feat, y)
        print("rows", n, "loss", want_loss, flush=True)
        assert np.array_equal(grad.view(np.uint32), want_grad.view(np.uint32)), n
        assert loss_sum / n == want_loss, n
    feat, y = _rows(limit + 1, 3)
    with pytest.raises(og.OgpuError, match="at most 16384"):
        _noalloc(lib, m, feat, y)
    m.free()
