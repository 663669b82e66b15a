"""ORACLE — test infrastructure only (never imported by the product).

FP64 numpy restatement of the operators of the learned preconditioner's
3-D U-Net, in the channels-last layout the HIP kernels use:

  conv3d_fwd(x[N,D,H,W,CI], w[CO,3,3,3,CI], b[CO], act) -> y[N,D,H,W,CO]
      3x3x3 cross-correlation, stride 1, zero padding 1, then
      act in {None, 'lrelu'} (leaky ReLU, negative slope 0.1);
  conv3d_bwd_data(dy, w)   -> dx   (adjoint of the convolution);
  conv3d_bwd_weight(x, dy) -> dw[CO,3,3,3,CI], db[CO];
  lrelu_bwd(dy, y)         -> dy * (y > 0 ? 1 : 0.1);
  avgpool2 / avgpool2_bwd, upsample2 / upsample2_bwd (nearest).

These have no counterpart in the reference (the learned preconditioner is
a project addition); they are pinned against torch.nn.functional.conv3d /
avg_pool3d / interpolate and autograd in float64 (tests/test_conv_oracle.py).
"""
import numpy as np

SLOPE = 0.1


def _pad(x):
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def conv3d_fwd(x, w, b=None, act=None):
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    N, D, H, W, CI = x.shape
    CO = w.shape[0]
    xp = _pad(x)
    y = np.zeros((N, D, H, W, CO))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                y += np.einsum("ndhwc,oc->ndhwo", xp[:, i:i + D, j:j + H, k:k + W], w[:, i, j, k])
    if b is not None:
        y += np.asarray(b, np.float64)
    if act == "lrelu":
        y = np.where(y > 0, y, SLOPE * y)
    return y


def conv3d_bwd_data(dy, w):
    dy = np.asarray(dy, np.float64)
    w = np.asarray(w, np.float64)
    N, D, H, W, CO = dy.shape
    CI = w.shape[-1]
    dx = np.zeros((N, D + 2, H + 2, W + 2, CI))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                dx[:, i:i + D, j:j + H, k:k + W] += np.einsum("ndhwo,oc->ndhwc", dy, w[:, i, j, k])
    return dx[:, 1:-1, 1:-1, 1:-1]


def conv3d_bwd_weight(x, dy):
    x = np.asarray(x, np.float64)
    dy = np.asarray(dy, np.float64)
    N, D, H, W, CI = x.shape
    CO = dy.shape[-1]
    xp = _pad(x)
    dw = np.zeros((CO, 3, 3, 3, CI))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                dw[:, i, j, k] = np.einsum("ndhwo,ndhwc->oc", dy, xp[:, i:i + D, j:j + H, k:k + W])
    return dw, dy.sum(axis=(0, 1, 2, 3))


def lrelu_bwd(dy, y):
    return np.where(np.asarray(y) > 0, dy, SLOPE * np.asarray(dy, np.float64))


def avgpool2(x):
    N, D, H, W, C = x.shape
    return np.asarray(x, np.float64).reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def avgpool2_bwd(dy):
    return np.repeat(np.repeat(np.repeat(np.asarray(dy, np.float64), 2, 1), 2, 2), 2, 3) / 8.0


def upsample2(x):
    return np.repeat(np.repeat(np.repeat(np.asarray(x, np.float64), 2, 1), 2, 2), 2, 3)


def upsample2_bwd(dy):
    N, D, H, W, C = dy.shape
    return np.asarray(dy, np.float64).reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))
