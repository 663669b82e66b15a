"""CPU tests pinning the U-Net operator oracle (oracle/conv_oracle.py)
against torch.nn.functional in float64: conv3d (zero padding 1) and its
autograd adjoints, leaky ReLU, avg_pool3d and nearest upsampling."""
import numpy as np
import torch
import torch.nn.functional as F

from oracle import conv_oracle as O


def _cf(a):  # channels-last numpy -> channels-first torch
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(a, -1, 1)))


def _cl(t):
    return np.moveaxis(t.detach().numpy(), 1, -1)


def test_conv3d_fwd_bwd_vs_torch():
    rng = np.random.default_rng(0)
    for (CI, CO, shape) in [(3, 5, (4, 5, 6)), (16, 8, (3, 4, 7))]:
        x = rng.standard_normal((2,) + shape + (CI,))
        w = rng.standard_normal((CO, 3, 3, 3, CI))
        b = rng.standard_normal(CO)
        xt = _cf(x).requires_grad_(True)
        wt = torch.from_numpy(np.ascontiguousarray(np.moveaxis(w, -1, 1))).requires_grad_(True)
        bt = torch.from_numpy(b).requires_grad_(True)
        yt = F.conv3d(xt, wt, bt, padding=1)
        at = F.leaky_relu(yt, O.SLOPE)
        y = O.conv3d_fwd(x, w, b, act="lrelu")
        assert np.abs(_cl(at) - y).max() < 1e-10
        assert np.abs(_cl(yt) - O.conv3d_fwd(x, w, b)).max() < 1e-10
        dy = rng.standard_normal(y.shape)
        at.backward(_cf(dy))
        dpre = O.lrelu_bwd(dy, y)
        assert np.abs(_cl(xt.grad) - O.conv3d_bwd_data(dpre, w)).max() < 1e-9
        dw, db = O.conv3d_bwd_weight(x, dpre)
        assert np.abs(np.moveaxis(wt.grad.numpy(), 1, -1) - dw).max() < 1e-9
        assert np.abs(bt.grad.numpy() - db).max() < 1e-9


def test_pool_and_upsample_vs_torch():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 6, 8, 3))
    xt = _cf(x).requires_grad_(True)
    yt = F.avg_pool3d(xt, 2)
    assert np.abs(_cl(yt) - O.avgpool2(x)).max() < 1e-12
    dy = rng.standard_normal(yt.shape)
    yt.backward(torch.from_numpy(dy))
    assert np.abs(_cl(xt.grad) - O.avgpool2_bwd(np.moveaxis(dy, 1, -1))).max() < 1e-12
    xt = _cf(x).requires_grad_(True)
    yt = F.interpolate(xt, scale_factor=2, mode="nearest")
    assert np.abs(_cl(yt) - O.upsample2(x)).max() < 1e-12
    dy = rng.standard_normal(yt.shape)
    yt.backward(torch.from_numpy(dy))
    assert np.abs(_cl(xt.grad) - O.upsample2_bwd(np.moveaxis(dy, 1, -1))).max() < 1e-12
