"""Element-wise accuracy of grid_push / grid_count on the GPU.

A norm-relative check (max|a - b| / max|b|) hides the error of small
outputs next to large ones.  Here every output element is held to its own
bar against the FP64 oracle on the same float32 inputs:

    |push_gpu - push_oracle| <= 1e-4 * push_oracle(|v|)        (element-wise)

where push_oracle(|v|) is the push of the absolute values, i.e. the sum of
|w| |v| over the splats that reach the element (the bounds used here have
non-negative weights).  The inputs mix magnitudes 1 and 1e-4 inside every
tile.  Non-finite values must propagate as the reference's float sums do:
NaN / Inf exactly where the oracle has them.

The tile push sums every box element in FP64 (LDS atomics), so the bar
holds however small an element is next to its tile's other values: a
further test puts regions of 1e-6-sized values beside values of 1, where
outputs reach 1e-15.  Tiny weights are evaluated from the coordinate
itself (the last spline piece, (h - |d|)^n / n!, with s = h - |d| taken as
an exact difference), so they too keep float's relative accuracy.
Reference semantics: nitorch/_C/src/impl/pushpull_common.cpp:1230-1285
(splat3d), :900 (out-of-bound tolerance).
"""
import numpy as np
import pytest
import torch

from oracle import pushpull as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
# float32 outputs: below float's normal range (1.2e-38) no relative bar holds
FLOAT_FLOOR = 1e-36


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _grid(rng, N, shape, amp=1.5, oob=0.05):
    axes = [np.arange(n, dtype=np.float64) for n in shape]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1)[None].repeat(N, 0)
    for n in range(N):
        for d in range(3):
            ph = rng.uniform(0, 2 * np.pi, 3)
            fr = rng.uniform(0.05, 0.3, 3)
            g[n, ..., d] += amp * np.sin(fr[0] * axes[0][:, None, None] + ph[0]) \
                * np.cos(fr[1] * axes[1][None, :, None] + ph[1]) * np.sin(fr[2] * axes[2][None, None, :] + ph[2])
    m = rng.random(g.shape[:-1]) < oob
    g[m] += rng.choice([-1, 1], (m.sum(), 1)) * np.array(shape) * (rng.random((m.sum(), 3)) < 0.4)
    return g.astype(np.float32)


def _mixed_values(rng, shape, tiny=None):
    """Random signs; magnitude 1 or 1e-4 per sample (every tile mixes both);
    with `tiny`, the small values of one x-half are `tiny` instead."""
    mag = np.where(rng.random(shape) < 0.5, 1.0, 1e-4)
    if tiny is not None:
        half = mag[..., : shape[-3] // 2, :, :]
        mag[..., : shape[-3] // 2, :, :] = np.where(half < 1, tiny, 1.0)
    return (mag * rng.choice([-1.0, 1.0], shape) * rng.uniform(0.5, 1.0, shape)).astype(np.float32)


def _check(out, ref, bar, what):
    out = out.detach().cpu().double().numpy()
    err = np.abs(out - ref)
    bad = err > 1e-4 * bar + FLOAT_FLOOR
    assert not bad.any(), (what, int(bad.sum()), float((err / np.maximum(bar, 1e-300))[bad].max()))


@pytest.mark.parametrize("order", [1, 2, 3, 5, 7])
def test_push_elementwise_mixed_magnitudes(order):
    rng = np.random.default_rng(900 + order)
    shape = (20, 35, 33)
    for bound in ("dct2", "dft", "replicate", "zero"):
        for extrap in (0, 1):
            g = _grid(rng, 2, shape, oob=0.0 if extrap else 0.05)
            y = _mixed_values(rng, (2, 3) + shape)
            from nitorch_amd import grid as G
            out = G.grid_push(T(y), T(g), list(shape), [bound], [order], extrap)
            g64, y64 = g.astype(np.float64), y.astype(np.float64)
            ref = O.grid_push(y64, g64, shape, bound, order, extrap, oob_f32=True)
            bar = O.grid_push(np.abs(y64), g64, shape, bound, order, extrap, oob_f32=True)
            _check(out, ref, bar, ("push", order, bound, extrap))
            cnt = G.grid_count(T(g), list(shape), [bound], [order], extrap)
            ref = O.grid_count(g64, shape, bound, order, extrap, oob_f32=True)
            _check(cnt, ref, ref, ("count", order, bound, extrap))


@pytest.mark.parametrize("order", [1, 3, 5])
def test_push_elementwise_tiny_regions(order):
    """Values of 1e-6 beside values of 1, same element-wise bar."""
    from nitorch_amd import grid as G
    rng = np.random.default_rng(970 + order)
    shape = (20, 35, 33)
    for bound in ("dct2", "zero"):
        for extrap in (0, 1):
            g = _grid(rng, 2, shape, oob=0.0 if extrap else 0.05)
            y = _mixed_values(rng, (2, 3) + shape, tiny=1e-6)
            out = G.grid_push(T(y), T(g), list(shape), [bound], [order], extrap)
            g64, y64 = g.astype(np.float64), y.astype(np.float64)
            ref = O.grid_push(y64, g64, shape, bound, order, extrap, oob_f32=True)
            bar = O.grid_push(np.abs(y64), g64, shape, bound, order, extrap, oob_f32=True)
            _check(out, ref, bar, ("push", order, bound, extrap))


@pytest.mark.parametrize("order", [0, 1, 3, 5, 7])
def test_push_propagates_nan_and_inf(order):
    """One NaN and one +Inf among finite values: the output is NaN / Inf at
    exactly the elements where the oracle's is, and meets the element-wise
    bar everywhere else."""
    from nitorch_amd import grid as G
    rng = np.random.default_rng(950 + order)
    shape = (16, 24, 32)
    g = _grid(rng, 1, shape, oob=0.0)
    y = rng.standard_normal((1, 2) + shape).astype(np.float32)
    y[0, 0, 5, 7, 9] = np.nan
    y[0, 1, 11, 3, 20] = np.inf
    out = G.grid_push(T(y), T(g), list(shape), ["dct2"], [order], 0).cpu().double().numpy()
    g64, y64 = g.astype(np.float64), y.astype(np.float64)
    ref = O.grid_push(y64, g64, shape, "dct2", order, 0, oob_f32=True)
    assert np.isnan(ref).any() and np.isinf(ref).any()
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    assert np.array_equal(np.isposinf(out), np.isposinf(ref))
    fin = np.isfinite(ref)
    bar = O.grid_push(np.where(np.isfinite(y64), np.abs(y64), 0.0), g64, shape, "dct2", order, 0, oob_f32=True)
    assert (np.abs(out[fin] - ref[fin]) <= 1e-4 * bar[fin] + FLOAT_FLOOR).all()
