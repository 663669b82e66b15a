"""GPU parity of the HIP push/pull kernels against the oracle.

Tolerances (norm-relative: max|a-b| / max|b|, written per test):
  * order-0 pulls: bit-exact;
  * pull / grad / grid gradients: 1e-5 in FP32;
  * push / count (atomic order varies): 1e-4 in FP32;
  * FP64 kernels against the FP64 oracle: 1e-10.
Every call goes through the C ABI (nitorch_amd.grid -> libnitorch_amd.so).
"""
import itertools

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


def identity(shape):
    return np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij"), -1)[None]


def make(rng, dim, N=2, C=3, sp=(7, 6, 5), spo=(5, 6, 4), scale=1.3, dtype=np.float32):
    sp, spo = sp[:dim], spo[:dim]
    x = rng.standard_normal((N, C) + sp).astype(dtype)
    g = identity(spo) * (np.array(sp) - 1) / np.maximum(np.array(spo) - 1, 1)
    g = (g + scale * rng.standard_normal((N,) + spo + (dim,))).astype(dtype)
    return x, g


def T(a, req=False):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).requires_grad_(req)


@pytest.fixture(scope="module")
def be():
    from nitorch_amd import grid as G
    return G


def _tol(dtype, kind):
    if dtype == np.float64:
        return 1e-10
    return {"gather": 1e-5, "scatter": 1e-4}[kind]


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("order", range(8))
def test_all_modes_vs_oracle(be, dim, order):
    rng = np.random.default_rng(100 * dim + order)
    for bound, extrap, dtype in itertools.product(BOUNDS, (0, 1), (np.float32, np.float64)):
        if dtype == np.float64 and bound not in ("dct2", "dst1", "zero"):
            continue
        x, g = make(rng, dim, dtype=dtype)
        B, I = [bound], [order]
        case = (dim, order, bound, extrap, dtype.__name__)
        # pull
        out = be.grid_pull(T(x), T(g), B, I, extrap)
        ref = O.grid_pull(x, g, bound, order, extrap)
        if order == 0:
            assert np.array_equal(out.cpu().numpy(), ref), case
        assert rel(out, ref) <= _tol(dtype, "gather"), ("pull",) + case
        # grad (spatial gradient)
        out = be.grid_grad(T(x), T(g), B, I, extrap)
        assert rel(out, O.grid_grad(x, g, bound, order, extrap)) <= _tol(dtype, "gather"), ("sgrad",) + case
        # push / count
        y = rng.standard_normal(x.shape[:2] + g.shape[1:-1]).astype(dtype)
        out = be.grid_push(T(y), T(g), list(x.shape[2:]), B, I, extrap)
        assert rel(out, O.grid_push(y, g, x.shape[2:], bound, order, extrap)) <= _tol(dtype, "scatter"), ("push",) + case
        out = be.grid_count(T(g), list(x.shape[2:]), B, I, extrap)
        assert rel(out, O.grid_count(g, x.shape[2:], bound, order, extrap)) <= _tol(dtype, "scatter"), ("count",) + case
        # pull backward (fused push + grid gradient)
        gi, gg = be.grid_pull_backward(T(y), T(x, True), T(g, True), B, I, extrap)
        ri, rg = O.grid_pull_backward(y, x, g, bound, order, extrap)
        assert rel(gi, ri) <= _tol(dtype, "scatter"), ("pull_bw_in",) + case
        assert rel(gg, rg) <= _tol(dtype, "gather"), ("pull_bw_grid",) + case
        # pull backward, one output at a time
        (gi1,) = be.grid_pull_backward(T(y), T(x, True), T(g, False), B, I, extrap)
        assert rel(gi1, ri) <= _tol(dtype, "scatter"), ("pull_bw_in_only",) + case
        (gg1,) = be.grid_pull_backward(T(y), T(x, False), T(g, True), B, I, extrap)
        assert rel(gg1, rg) <= _tol(dtype, "gather"), ("pull_bw_grid_only",) + case
        # push backward
        gi, gg = be.grid_push_backward(T(x), T(y, True), T(g, True), B, I, extrap)
        ri, rg = O.grid_push_backward(x, y, g, bound, order, extrap)
        assert rel(gi, ri) <= _tol(dtype, "gather"), ("push_bw_in",) + case
        assert rel(gg, rg) <= _tol(dtype, "gather"), ("push_bw_grid",) + case
        # count backward
        gc = rng.standard_normal((x.shape[0], 1) + x.shape[2:]).astype(dtype)
        out = be.grid_count_backward(T(gc), T(g, True), B, I, extrap)
        assert rel(out, O.grid_count_backward(gc, g, bound, order, extrap)) <= _tol(dtype, "gather"), ("count_bw",) + case
        # grad backward (input: push of the spatial-gradient adjoint; grid: 2nd derivatives)
        gs = rng.standard_normal(x.shape[:2] + g.shape[1:-1] + (dim,)).astype(dtype)
        gi, gg = be.grid_grad_backward(T(gs), T(x, True), T(g, True), B, I, extrap)
        ri, rg = O.grid_grad_backward(gs, x, g, bound, order, extrap)
        assert rel(gi, ri) <= _tol(dtype, "scatter"), ("grad_bw_in",) + case
        assert rel(gg, rg) <= _tol(dtype, "gather"), ("grad_bw_grid",) + case


def test_anisotropic_orders(be):
    rng = np.random.default_rng(7)
    for orders in ([1, 3, 3], [0, 2, 5], [3, 1, 0], [7, 4, 2]):
        for bound in (["dct2"], ["dft", "zero", "dst2"]):
            x, g = make(rng, 3)
            out = be.grid_pull(T(x), T(g), bound, orders, 0)
            ref = O.grid_pull(x, g, bound, orders, 0)
            assert rel(out, ref) <= 1e-5, (orders, bound)
            y = rng.standard_normal(x.shape[:2] + g.shape[1:-1]).astype(np.float32)
            out = be.grid_push(T(y), T(g), list(x.shape[2:]), bound, orders, 0)
            assert rel(out, O.grid_push(y, g, x.shape[2:], bound, orders, 0)) <= 1e-4


def test_strided_inputs(be):
    """Non-contiguous input, grid and upstream gradients (any strides)."""
    rng = np.random.default_rng(8)
    x, g = make(rng, 3, C=4)
    xt = T(x).permute(0, 1, 4, 3, 2).contiguous().permute(0, 1, 4, 3, 2)      # channel/spatial strides
    gt = T(np.moveaxis(g, -1, 1).copy()).movedim(1, -1)                          # grid last-dim stride > 1
    for order in (0, 1, 3):
        out = be.grid_pull(xt, gt, ["dct2"], [order], 0)
        assert rel(out, O.grid_pull(x, g, "dct2", order, 0)) <= 1e-5
        gs = T(rng.standard_normal((2, 4, 3) + g.shape[1:-1]).astype(np.float32)).movedim(2, -1)
        gi, gg = be.grid_grad_backward(gs, xt.detach().requires_grad_(), gt.detach().requires_grad_(),
                                       ["dct2"], [order], 0)
        ri, rg = O.grid_grad_backward(gs.contiguous().cpu().numpy(), x, g, "dct2", order, 0)
        assert rel(gi, ri) <= 1e-4 and rel(gg, rg) <= 1e-5


def test_public_api_broadcast_and_autograd():
    from nitorch_amd import spatial
    rng = np.random.default_rng(9)
    x = torch.from_numpy(rng.standard_normal((3, 8, 9, 7))).to(DEV)             # [C, *spatial]
    g = spatial.identity_grid((6, 7, 5), dtype=torch.float64, device=DEV)
    g = (g + torch.from_numpy(rng.standard_normal((2, 6, 7, 5, 3))).to(DEV)).requires_grad_()
    x.requires_grad_()
    y = spatial.grid_pull(x, g, interpolation=3, bound="dft", extrapolate=True)
    assert y.shape == (2, 3, 6, 7, 5)
    ref = O.grid_pull(np.broadcast_to(x.detach().cpu().numpy(), (2, 3, 8, 9, 7)), g.detach().cpu().numpy(), "dft", 3, 1)
    assert rel(y, ref) < 1e-10
    w = torch.randn_like(y)
    (y * w).sum().backward()
    ri, rg = O.grid_pull_backward(w.cpu().numpy(), np.broadcast_to(x.detach().cpu().numpy(), (2, 3, 8, 9, 7)),
                                  g.detach().cpu().numpy(), "dft", 3, 1)
    assert rel(x.grad, ri.sum(0)) < 1e-10
    assert rel(g.grad, rg) < 1e-10
    # label maps
    lab = torch.from_numpy(rng.integers(0, 4, (8, 9, 7))).to(DEV)
    out = spatial.grid_pull(lab, spatial.identity_grid((8, 9, 7), dtype=torch.float32, device=DEV), 1, "dct2")
    assert torch.equal(out, lab)


@pytest.mark.parametrize("fn", ["pull", "push", "count", "grad"])
def test_gradcheck_like_reference(fn):
    """torch.autograd.gradcheck in float64 as nitorch/tests/test_gradcheck_spatial.py:69-110
    (spatial size 3, identity + N(0,1) grid, extrapolate=True)."""
    from nitorch_amd import spatial
    torch.manual_seed(0)
    for dim in (1, 2, 3):
        for order in (0, 1, 2, 3, 5):
            for bound in ("dct2", "dft", "zero", "dst1"):
                shape = (3,) * dim
                g = spatial.identity_grid(shape, dtype=torch.float64, device=DEV)
                g = (g + torch.randn_like(g)).requires_grad_()
                x = torch.randn((1,) + shape, dtype=torch.float64, device=DEV, requires_grad=True)
                if fn == "pull":
                    f = lambda x_, g_: spatial.grid_pull(x_, g_, order, bound, True)
                    args = (x, g)
                elif fn == "push":
                    f = lambda x_, g_: spatial.grid_push(x_, g_, shape, order, bound, True)
                    args = (x, g)
                elif fn == "count":
                    f = lambda g_: spatial.grid_count(g_, shape, order, bound, True)
                    args = (g,)
                else:
                    f = lambda x_, g_: spatial.grid_grad(x_, g_, order, bound, True)
                    args = (x, g)
                assert torch.autograd.gradcheck(f, args, rtol=1.0, raise_exception=True), (fn, dim, order, bound)


def test_config1_full_size():
    """Plain pull 1x1x96^3, order 1, dct2 (CPU-runnable configuration) against the oracle."""
    from nitorch_amd import spatial
    from nitorch_amd.bench_data import config1_inputs
    x, g = config1_inputs()
    out = spatial.grid_pull(x.to(DEV), g.to(DEV), 1, "dct2", False)
    ref = O.grid_pull(x.numpy(), g.numpy(), "dct2", 1, 0)
    assert rel(out, ref) <= 1e-5


def test_config3_directional_finite_differences():
    """backward_128 recipe (order 3, dft, upstream gradient N(0,1)) generated
    at 24^3, in float64: the grid and input gradients of L = <w, pull(x, g)>
    against central differences along random directions (step 1e-4), and a
    per-entry check on a few voxels.  extrapolate=True, so that no sample
    crosses the out-of-bound cut within the step (the loss is smooth)."""
    from nitorch_amd import grid as G
    from nitorch_amd.bench_data import config3_inputs
    x, g, w = (t.double().to(DEV).contiguous() for t in config3_inputs(batch=1, shape=(24, 24, 24)))
    B, O, E = ["dft"], [3], 1

    def L(x_, g_):
        return float((G.grid_pull(x_, g_, B, O, E) * w).sum())

    # outputs selected by requires_grad, as the reference's deque
    gi, gg = G.grid_pull_backward(w, x.clone().requires_grad_(), g.clone().requires_grad_(), B, O, E)
    rng = np.random.default_rng(0)
    h = 1e-4
    for _ in range(4):
        vg = torch.from_numpy(rng.standard_normal(g.shape)).to(DEV)
        vx = torch.from_numpy(rng.standard_normal(x.shape)).to(DEV)
        fd_g = (L(x, g + h * vg) - L(x, g - h * vg)) / (2 * h)
        fd_x = (L(x + h * vx, g) - L(x - h * vx, g)) / (2 * h)
        assert abs(fd_g - float((gg * vg).sum())) <= 1e-6 * abs(fd_g) + 1e-8
        assert abs(fd_x - float((gi * vx).sum())) <= 1e-8 * abs(fd_x) + 1e-8
    for (i, j, k, d) in [(0, 0, 0, 0), (5, 11, 23, 1), (12, 3, 7, 2), (23, 23, 23, 0)]:
        e = torch.zeros_like(g)
        e[0, i, j, k, d] = 1
        fd = (L(x, g + h * e) - L(x, g - h * e)) / (2 * h)
        assert abs(fd - float(gg[0, i, j, k, d])) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("prefilter", [False, True])
def test_label_map_pull_vs_oracle(prefilter):
    """Integer inputs are pulled as one soft mask per label and the arg-max
    label kept (nitorch/spatial/_grid.py:187-198), here against the same
    procedure on the oracle's pull.  With `prefilter`, the reference assigns
    the prefiltered mask to `input` (in place), so later labels are compared
    with coefficients: reproduced."""
    from nitorch_amd import spatial
    from oracle import coeff as CF
    rng = np.random.default_rng(17)
    lab = rng.integers(0, 5, (2, 1, 9, 8, 7))
    grid = (O_identity((9, 8, 7))[None] + rng.uniform(-1.5, 1.5, (2, 9, 8, 7, 3)))
    order = 3 if prefilter else 1
    out = spatial.grid_pull(torch.from_numpy(lab).to(DEV), torch.from_numpy(grid).to(DEV), interpolation=order,
                            bound="dct2", extrapolate=False, prefilter=prefilter)
    ref = np.zeros(lab.shape, dtype=lab.dtype)
    pmax = np.zeros(lab.shape)
    inp = lab
    for label in np.unique(lab):
        soft = (inp == label).astype(np.float64)
        if prefilter:
            soft = CF.spline_coeff_nd(soft, "dct2", order, dim=3)
            inp = soft
        s = O.grid_pull(soft, grid, "dct2", order, 0)
        ref[s > pmax] = label
        pmax = np.maximum(pmax, s)
    assert out.dtype == torch.from_numpy(lab).dtype
    assert np.array_equal(out.cpu().numpy(), ref)


def O_identity(shape):
    return np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij"), -1)


def test_int64_offsets():
    """Tensors whose largest element offset does not fit 32 bits take the
    64-bit-offset direct kernel (the reference's int64 dispatch,
    pushpull_common.cpp:2506-2519).  Here the source and the pushed values are
    two small volumes 2^31 elements apart in one 8 GiB buffer (batch stride
    2^31): pull, grid_grad and push must equal the oracle -- order 0
    bit-exactly -- as on contiguous copies, which take the 32-bit path."""
    from nitorch_amd import grid as G
    rng = np.random.default_rng(77)
    shape = (9, 10, 11)
    V, S = 990, 2 ** 31
    big = torch.zeros(S + 4 * 65536, device=DEV)
    strides = (S, V, 110, 11, 1)
    x = rng.standard_normal((2, 2) + shape).astype(np.float32)
    y = rng.standard_normal((2, 2) + shape).astype(np.float32)
    src = big.as_strided((2, 2) + shape, strides, 0)
    val = big.as_strided((2, 2) + shape, strides, 65536)
    src.copy_(torch.from_numpy(x))
    val.copy_(torch.from_numpy(y))
    g = identity(shape).repeat(2, 0) + rng.uniform(-1.5, 1.5, (2,) + shape + (3,))
    g = g.astype(np.float32)
    g64, x64, y64 = g.astype(np.float64), x.astype(np.float64), y.astype(np.float64)
    try:
        for order in (0, 1, 3):
            for bound in ("dct2", "dft", "zero"):
                case = (order, bound)
                out = G.grid_pull(src, T(g), [bound], [order], 0).cpu().double().numpy()
                ref = O.grid_pull(x64, g64, bound, order, 0, oob_f32=True)
                if order == 0:
                    assert np.array_equal(out, ref), ("pull",) + case
                assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max(), ("pull",) + case
                out32 = G.grid_pull(T(x), T(g), [bound], [order], 0).cpu().double().numpy()
                assert np.abs(out - out32).max() <= 1e-5 * np.abs(ref).max(), ("pull 32 vs 64",) + case
                if order > 0:
                    out = G.grid_grad(src, T(g), [bound], [order], 0).cpu().double().numpy()
                    ref = O.grid_grad(x64, g64, bound, order, 0, oob_f32=True)
                    assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max(), ("grad",) + case
                out = G.grid_push(val, T(g), list(shape), [bound], [order], 0).cpu().double().numpy()
                ref = O.grid_push(y64, g64, shape, bound, order, 0, oob_f32=True)
                bar = O.grid_push(np.abs(y64), g64, shape, bound, order, 0, oob_f32=True)
                assert (np.abs(out - ref) <= 1e-4 * bar + 1e-36).all(), ("push",) + case
    finally:
        del src, val, big
        torch.cuda.empty_cache()
