"""GPU parity of the LDS-tiled pull / push / count path (3-D, float32,
isotropic orders) against the oracle.

Smooth deformations keep every work-group's tap box inside LDS (tile path);
tiles with scattered coordinates take the direct per-sample kernel; the
mixed case puts both kinds of tile in one call.  Shapes are not multiples
of the 4 x 8 x 16 tile, channel counts exercise partial channel chunks, and
a share of the samples is moved out of bounds.

Tolerances (norm-relative, max|a-b| / max|b|): pull 1e-5 (order 0:
bit-exact), push / count 1e-4 (summation order differs).
"""
import numpy as np
import pytest
import torch

from oracle import pushpull as O

pytestmark = pytest.mark.gpu

BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]
DEV = "cuda"


def rel(a, b):
    a = a.detach().cpu().double().numpy() if torch.is_tensor(a) else np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max()) / max(float(np.abs(b).max()), 1e-30)


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def smooth_grid(rng, N, out_shape, in_shape, amp=1.2, oob=0.05, scatter=0.0):
    """identity (rescaled to the input) + a smooth displacement of amplitude
    `amp` voxels; a fraction `oob` of points pushed out of bounds and a
    fraction `scatter` of 4x8x16 tiles given random coordinates."""
    axes = [np.linspace(0, i - 1, o) for o, i in zip(out_shape, in_shape)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1)[None].repeat(N, 0)
    for n in range(N):
        for d in range(3):
            ph = rng.uniform(0, 2 * np.pi, 3)
            fr = rng.uniform(0.05, 0.25, 3)
            disp = amp * np.sin(fr[0] * axes[0][:, None, None] + ph[0]) \
                * np.cos(fr[1] * axes[1][None, :, None] + ph[1]) \
                * np.sin(fr[2] * axes[2][None, None, :] + ph[2])
            g[n, ..., d] += disp
    if oob:
        m = rng.random(g.shape[:-1]) < oob
        ax = rng.integers(0, 3, m.sum())
        sgn = rng.choice([-1, 1], m.sum())
        pts = g[m]
        pts[np.arange(len(ax)), ax] += sgn * np.array(in_shape)[ax]
        g[m] = pts
    if scatter:
        tiles = [(n, x, y, z) for n in range(N) for x in range(0, out_shape[0], 4)
                 for y in range(0, out_shape[1], 8) for z in range(0, out_shape[2], 16)]
        pick = rng.random(len(tiles)) < scatter
        for (n, x, y, z), p in zip(tiles, pick):
            if p:
                blk = g[n, x:x + 4, y:y + 8, z:z + 16]
                blk[...] = rng.uniform(-0.5, 1.0, blk.shape) * (np.array(in_shape) - 0.5)
    return g.astype(np.float32)


@pytest.fixture(scope="module")
def be():
    from nitorch_amd import grid as G
    return G


@pytest.mark.parametrize("order", range(8))
def test_tile_path_vs_oracle(be, order):
    rng = np.random.default_rng(500 + order)
    in_shape, out_shape = (19, 21, 37), (18, 23, 35)
    for k, C in enumerate((1, 2, 3, 8)):
        bound = BOUNDS[(order + k) % len(BOUNDS)]
        for extrap in (0, 1):
            case = (order, C, bound, extrap)
            g = smooth_grid(rng, 2, out_shape, in_shape, oob=0.0 if extrap else 0.05)
            x = rng.standard_normal((2, C) + in_shape).astype(np.float32)
            out = be.grid_pull(T(x), T(g), [bound], [order], extrap)
            ref = O.grid_pull(x, g, bound, order, extrap)
            if order == 0:
                assert np.array_equal(out.cpu().numpy(), ref), ("pull",) + case
            assert rel(out, ref) <= 1e-5, ("pull",) + case
            y = rng.standard_normal((2, C) + out_shape).astype(np.float32)
            out = be.grid_push(T(y), T(g), list(in_shape), [bound], [order], extrap)
            assert rel(out, O.grid_push(y, g, in_shape, bound, order, extrap)) <= 1e-4, ("push",) + case
            out = be.grid_count(T(g), list(in_shape), [bound], [order], extrap)
            assert rel(out, O.grid_count(g, in_shape, bound, order, extrap)) <= 1e-4, ("count",) + case


@pytest.mark.parametrize("order", [1, 3, 5])
def test_tile_and_fallback_mixed(be, order):
    """Smooth and scattered tiles in one call: both kernels run, each tile
    exactly once (no tile is dropped or done twice)."""
    rng = np.random.default_rng(600 + order)
    in_shape, out_shape = (24, 20, 33), (24, 20, 33)
    for bound in ("dct2", "dft", "zero"):
        g = smooth_grid(rng, 1, out_shape, in_shape, scatter=0.3)
        x = rng.standard_normal((1, 5) + in_shape).astype(np.float32)
        out = be.grid_pull(T(x), T(g), [bound], [order], 0)
        assert rel(out, O.grid_pull(x, g, bound, order, 0)) <= 1e-5, ("pull", bound)
        y = rng.standard_normal((1, 5) + out_shape).astype(np.float32)
        out = be.grid_push(T(y), T(g), list(in_shape), [bound], [order], 0)
        assert rel(out, O.grid_push(y, g, in_shape, bound, order, 0)) <= 1e-4, ("push", bound)
        out = be.grid_count(T(g), list(in_shape), [bound], [order], 0)
        assert rel(out, O.grid_count(g, in_shape, bound, order, 0)) <= 1e-4, ("count", bound)


def test_tile_strided_input(be):
    """Non-contiguous source / target / grid views through the tile path."""
    rng = np.random.default_rng(700)
    in_shape, out_shape = (20, 18, 40), (17, 19, 33)
    g = smooth_grid(rng, 2, out_shape, in_shape)
    x = rng.standard_normal((2, 3) + in_shape).astype(np.float32)
    xt = T(np.ascontiguousarray(np.moveaxis(x, 1, -1))).movedim(-1, 1)       # channel-last storage
    gt = T(np.ascontiguousarray(np.moveaxis(g, -1, 1))).movedim(1, -1)       # coordinate-first storage
    out = be.grid_pull(xt, gt, ["dct2"], [3], 0)
    assert rel(out, O.grid_pull(x, g, "dct2", 3, 0)) <= 1e-5
    y = rng.standard_normal((2, 3) + out_shape).astype(np.float32)
    yt = T(np.ascontiguousarray(np.moveaxis(y, 1, -1))).movedim(-1, 1)
    out = be.grid_push(yt, gt, list(in_shape), ["dft"], [2], 0)
    assert rel(out, O.grid_push(y, g, in_shape, "dft", 2, 0)) <= 1e-4
