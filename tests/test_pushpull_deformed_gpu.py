"""GPU parity of pull / push / count / grad on strongly deformed grids, against the
FP64 oracle: rotations with zoom (affine lattices) and smooth displacements
with a large Jacobian, and samples collapsed at random into a small
region.  Their tiles do not fit the static LDS box.  In calls of at least
2^20 samples the listed tiles' samples are binned by first tap (a few bins
of the collapsed grid hold thousands of samples, processed in batches); in
smaller calls each listed tile is cut into sub-tiles with boxes of their
own.  Both routes are tested.

Tolerances: pull and grad 1e-5 norm-relative (pull at order 0: bit-exact); push / count
element-wise |gpu - oracle| <= 1e-4 * oracle(|v|) (+ 1e-36, below float's
normal range).  The affine lattices put
float32 coordinates exactly on the out-of-bound edges; the oracle decides
those in float, as the reference does for float data (`oob_f32`).
"""
import math

import numpy as np
import pytest
import torch

from oracle import pushpull as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def affine(shape, angle, axis, zoom, shift):
    c, s = math.cos(math.radians(angle)), math.sin(math.radians(angle))
    R = np.eye(3)
    i, j = [k for k in range(3) if k != axis]
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    ctr = (np.array(shape) - 1) / 2
    ijk = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij"), -1) - ctr
    return (ijk @ (zoom * R).T + ctr + np.array(shift))[None]


def warped(rng, shape, amp, freq):
    axes = [np.arange(n, dtype=np.float64) for n in shape]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1)[None]
    for d in range(3):
        ph = rng.uniform(0, 2 * np.pi, 3)
        g[0, ..., d] += amp * np.sin(freq * axes[0][:, None, None] + ph[0]) \
            * np.cos(freq * axes[1][None, :, None] + ph[1]) * np.sin(freq * axes[2][None, None, :] + ph[2])
    return g


def grids(rng, shape):
    yield "rot20z_zoom1.15", affine(shape, 20, 2, 1.15, (0.3, -0.7, 0.2))
    yield "rot35x_zoom0.8", affine(shape, 35, 0, 0.8, (1.1, 0.4, -0.5))
    yield "warp_amp6", warped(rng, shape, 6.0, 0.15)
    # every sample scattered at random inside a 6^3 region: the tiles are
    # too spread for the tile box and a few bins hold thousands of samples
    yield "collapse6", (10 + 6 * rng.random((1,) + shape + (3,)))
    # the same with a tenth of the samples 40 voxels out along one axis:
    # out of bounds, or (extrapolate=1) in bounds but beyond every bin,
    # so they run the per-sample routine
    g = 10 + 6 * rng.random((1,) + shape + (3,))
    far = rng.random(shape) < 0.1
    ax = rng.integers(0, 3, shape)
    for d in range(3):
        g[0][far & (ax == d), d] += rng.choice([-40.0, 40.0], int((far & (ax == d)).sum()))
    yield "collapse6_far", g


@pytest.mark.parametrize("order", [0, 1, 3, 5])
def test_deformed_pull_push_count(order):
    from nitorch_amd import grid as G
    rng = np.random.default_rng(1100 + order)
    shape = (37, 41, 45)
    for name, g in grids(rng, shape):
        g = g.astype(np.float32)
        g64 = g.astype(np.float64)
        for bound in ("dct2", "dft", "zero"):
            for extrap in (0, 1):
                case = (name, order, bound, extrap)
                x = rng.standard_normal((1, 3) + shape).astype(np.float32)
                out = G.grid_pull(T(x), T(g), [bound], [order], extrap).cpu().double().numpy()
                ref = O.grid_pull(x.astype(np.float64), g64, bound, order, extrap, oob_f32=True)
                if order == 0:
                    assert np.array_equal(out, ref), ("pull",) + case
                assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max(), ("pull",) + case
                if order > 0:
                    out = G.grid_grad(T(x), T(g), [bound], [order], extrap).cpu().double().numpy()
                    ref = O.grid_grad(x.astype(np.float64), g64, bound, order, extrap, oob_f32=True)
                    assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max(), ("grad",) + case
                y = rng.standard_normal((1, 3) + shape).astype(np.float32)
                y64 = y.astype(np.float64)
                out = G.grid_push(T(y), T(g), list(shape), [bound], [order], extrap).cpu().double().numpy()
                ref = O.grid_push(y64, g64, shape, bound, order, extrap, oob_f32=True)
                bar = O.grid_push(np.abs(y64), g64, shape, bound, order, extrap, oob_f32=True)
                assert (np.abs(out - ref) <= 1e-4 * bar + 1e-36).all(), ("push",) + case
                out = G.grid_count(T(g), list(shape), [bound], [order], extrap).cpu().double().numpy()
                ref = O.grid_count(g64, shape, bound, order, extrap, oob_f32=True)
                assert (np.abs(out - ref) <= 1e-4 * ref + 1e-36).all(), ("count",) + case


@pytest.mark.parametrize("order", [3])
def test_deformed_binned_pass(order):
    """The same checks at 100 x 104 x 104 (>= 2^20 samples: the binned second
    pass, grid_grad's included)."""
    from nitorch_amd import grid as G
    rng = np.random.default_rng(1200 + order)
    shape = (100, 104, 104)
    for name, g in grids(rng, shape):
        if name in ("rot35x_zoom0.8", "collapse6_far"):
            continue
        g = g.astype(np.float32)
        g64 = g.astype(np.float64)
        for bound in ("dct2", "dft"):
            extrap = 1 if bound == "dft" else 0
            case = (name, order, bound, extrap)
            x = rng.standard_normal((1, 2) + shape).astype(np.float32)
            out = G.grid_pull(T(x), T(g), [bound], [order], extrap).cpu().double().numpy()
            ref = O.grid_pull(x.astype(np.float64), g64, bound, order, extrap, oob_f32=True)
            assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max(), ("pull",) + case
            out = G.grid_grad(T(x), T(g), [bound], [order], extrap).cpu().double().numpy()
            ref = O.grid_grad(x.astype(np.float64), g64, bound, order, extrap, oob_f32=True)
            assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max(), ("grad",) + case
            y = rng.standard_normal((1, 2) + shape).astype(np.float32)
            y64 = y.astype(np.float64)
            out = G.grid_push(T(y), T(g), list(shape), [bound], [order], extrap).cpu().double().numpy()
            ref = O.grid_push(y64, g64, shape, bound, order, extrap, oob_f32=True)
            bar = O.grid_push(np.abs(y64), g64, shape, bound, order, extrap, oob_f32=True)
            assert (np.abs(out - ref) <= 1e-4 * bar + 1e-36).all(), ("push",) + case
            out = G.grid_count(T(g), list(shape), [bound], [order], extrap).cpu().double().numpy()
            ref = O.grid_count(g64, shape, bound, order, extrap, oob_f32=True)
            assert (np.abs(out - ref) <= 1e-4 * ref + 1e-36).all(), ("count",) + case
