"""ni_grid_push_noalloc (the push entry point that allocates nothing: its
tile list lives in a device-resident array) against ni_grid_push, through
the C ABI: two calls in a row, the second on a smaller input, each to the
push parity bar (element-wise |a - b| <= 1e-4 * push(|v|)), and the error
for an input beyond the scratch capacity."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _fn():
    from nitorch_amd import _lib as L
    f = L.lib().ni_grid_push_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [L._I32, L._I32, L._PI32, L._PI32, L._I32, L._I64, L._I64, L._PI64, L._PI64,
                  L._P, L._PI64, L._P, L._PI64, L._P, L._PI64, L._P]
    return f


def _push_noalloc(y, g, shape, bound, order):
    from nitorch_amd import _lib as L
    from nitorch_amd import grid as G
    N, C = y.shape[:2]
    out = torch.empty((N, C) + tuple(shape), dtype=y.dtype, device=y.device)
    b, o = G._codes([bound], [order])
    rc = _fn()(L.dtype_code(y.dtype), 3, b, o, 0, N, C, L.i64arr(shape), L.i64arr(g.shape[1:-1]),
               L.ptr(y), L.strides(y), L.ptr(g), L.strides(g), L.ptr(out), L.strides(out), L.stream_of(g))
    L.check(rc, "grid_push_noalloc")
    return out


def _grid(rng, N, shape, amp):
    axes = [np.arange(n, dtype=np.float64) for n in shape]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1)[None].repeat(N, 0)
    g += amp * rng.standard_normal(g.shape) * 0.3
    # a few scattered samples: listed tiles
    m = rng.random(g.shape[:-1]) < 0.01
    g[m] = rng.uniform(0, 1, (m.sum(), 3)) * (np.array(shape) - 1)
    return torch.from_numpy(g.astype(np.float32)).to(DEV)


@pytest.mark.parametrize("order", [1, 3, 5])
def test_noalloc_push_matches_push(order):
    from nitorch_amd import grid as G
    rng = np.random.default_rng(1300 + order)
    for shape, N, C in (((40, 36, 50), 2, 5), ((17, 21, 19), 1, 3)):   # second call smaller
        g = _grid(rng, N, shape, 1.0)
        y = torch.from_numpy(rng.standard_normal((N, C) + shape).astype(np.float32)).to(DEV)
        ref = G.grid_push(y, g, list(shape), ["dct2"], [order], 0).double()
        bar = G.grid_push(y.abs(), g, list(shape), ["dct2"], [order], 0).double()
        out = _push_noalloc(y, g, shape, "dct2", order).double()
        torch.cuda.synchronize()
        assert bool(((out - ref).abs() <= 1e-4 * bar + 1e-30).all()), (order, shape)


def test_noalloc_push_rejects_oversize():
    from nitorch_amd import _lib as L
    from nitorch_amd import grid as G
    y = torch.zeros((1, 1, 2, 2, 2), device=DEV)
    g = torch.zeros((1, 2, 2, 2, 3), device=DEV)
    b, o = G._codes(["dct2"], [1])
    # 4096 x 8192 x 16 samples = 2^20 tiles of 4 x 8 x 16: refused before any launch
    rc = _fn()(L.dtype_code(y.dtype), 3, b, o, 0, 1, 1, L.i64arr((2, 2, 2)), L.i64arr((4096, 8192, 16)),
               L.ptr(y), L.strides(y), L.ptr(g), L.strides(g), L.ptr(y), L.strides(y), L.stream_of(g))
    assert rc != 0 and "scratch capacity" in L.last_error()
