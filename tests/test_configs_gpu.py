"""The benchmark workloads at their full sizes, checked on the GPU.

* pushpull_192 (4x8x192^3, grid = identity + smoothed displacement, 5% out
  of bounds), as generated and with the displacement rescaled to its
  nominal std of 8 voxels: orders {0, 1, 3, 5} x bounds {dct2, dft,
  replicate, zero} for pull, push, count and grad, each launched on the
  full-size tensors and compared with the FP64 oracle on a slab (batch 0,
  the first x-slices of the grid; push / count with every sample outside
  the slab given a zero value / moved out of bounds, so that the full-size
  launch computes exactly the slab's splats).  Full-size size-independent
  checks: adjointness <pull x, y> = <x, push y> and count mass (dct2, dft,
  replicate), summed in FP64.
* backward_128 (2x3x128x160x128, order 3, dft): both gradients of the
  pull, full size, against the oracle on a slab.
* fmg_160 (160x192x160x3): the plain FMG solve (both smoothers) and the
  learned-preconditioner PCG to 1e-4, residual recomputed by the FP64
  oracle operator.

Tolerances: pull / grad 1e-5 norm-relative over the slab (order 0:
bit-exact); push / count element-wise |gpu - oracle| <= 1e-4 * oracle(|v|)
(+ 1e-36, below float's normal range); the oracle decides out-of-bound
samples in float, as the reference does for float data;
adjointness / mass 1e-5 relative; solves <= 1.2e-4 FP64 residual (the
solver stops at 1e-4 measured in FP32).
"""
import numpy as np
import pytest
import torch

from oracle import pushpull as O

pytestmark = pytest.mark.gpu
DEV = "cuda"
ORDERS = (0, 1, 3, 5)
BOUNDS = ("dct2", "dft", "replicate", "zero")
SLAB = {0: 8, 1: 8, 3: 4, 5: 1}     # x-slices of batch 0 compared with the oracle


def _np(t):
    return t.detach().cpu().double().numpy()


def _normrel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module", params=["literal", "rescaled"])
def c2(request):
    from nitorch_amd.bench_data import config2_inputs
    x, g = config2_inputs(rescale=request.param == "rescaled")
    gen = torch.Generator().manual_seed(22)
    y = torch.rand(x.shape, generator=gen)
    return request.param, x, g, y


def test_c2_slab_sweep(c2):
    from nitorch_amd import grid as G
    name, x, g, y = c2
    xd, gd, yd = x.to(DEV), g.to(DEV), y.to(DEV)
    shape = list(x.shape[2:])
    x0 = x[:1].double().numpy()
    for order in ORDERS:
        k = SLAB[order]
        gs = g[:1, :k].double().numpy()
        ys = y[:1, :, :k].double().numpy()
        # push / count inputs restricted to the slab
        ym = torch.zeros_like(yd)
        ym[:1, :, :k] = yd[:1, :, :k]
        gm = torch.full_like(gd, -1e4)
        gm[:1, :k] = gd[:1, :k]
        for bound in BOUNDS:
            case = (name, order, bound)
            print("c2", case, flush=True)
            out = G.grid_pull(xd, gd, [bound], [order], 0)
            ref = O.grid_pull(x0, gs, bound, order, 0, oob_f32=True)
            got = _np(out[:1, :, :k])
            if order == 0:
                assert np.array_equal(got, ref), ("pull",) + case
            assert _normrel(got, ref) <= 1e-5, ("pull",) + case
            out = G.grid_grad(xd, gd, [bound], [order], 0)
            ref = O.grid_grad(x0, gs, bound, order, 0, oob_f32=True)
            assert _normrel(_np(out[:1, :, :k]), ref) <= 1e-5, ("grad",) + case
            out = _np(G.grid_push(ym, gd, shape, [bound], [order], 0)[:1])
            ref = O.grid_push(ys, gs, shape, bound, order, 0, oob_f32=True)
            bar = O.grid_push(np.abs(ys), gs, shape, bound, order, 0, oob_f32=True)
            assert (np.abs(out - ref) <= 1e-4 * bar + 1e-36).all(), ("push",) + case
            out = _np(G.grid_count(gm, shape, [bound], [order], 0)[:1])
            ref = O.grid_count(gs, shape, bound, order, 0, oob_f32=True)
            assert (np.abs(out - ref) <= 1e-4 * ref + 1e-36).all(), ("count",) + case


def test_c2_full_size_adjoint_and_mass(c2):
    from nitorch_amd import grid as G
    name, x, g, y = c2
    xd, gd, yd = x.to(DEV), g.to(DEV), y.to(DEV)
    shape = list(x.shape[2:])
    tol = 0.05
    inb = ((gd >= -tol) & (gd < torch.tensor(shape, device=DEV, dtype=gd.dtype) - 1 + tol)).all(-1)
    n_in = inb.sum(dtype=torch.float64).item()
    for order in ORDERS:
        for bound in BOUNDS:
            case = (name, order, bound)
            lhs = (G.grid_pull(xd, gd, [bound], [order], 0).double() * yd.double()).sum().item()
            rhs = (xd.double() * G.grid_push(yd, gd, shape, [bound], [order], 0).double()).sum().item()
            assert abs(lhs - rhs) <= 1e-5 * abs(lhs), ("adjoint",) + case
            if bound != "zero":
                mass = G.grid_count(gd, shape, [bound], [order], 0).double().sum().item()
                assert abs(mass - n_in) <= 1e-5 * n_in, ("mass",) + case


def test_c3_backward_full_size():
    from nitorch_amd import grid as G
    from nitorch_amd.bench_data import config3_inputs
    x, g, d = config3_inputs()
    xd, gd, dd = x.to(DEV), g.to(DEV), d.to(DEV)
    k = 4
    x0, gs, ds = x[:1].double().numpy(), g[:1, :k].double().numpy(), d[:1, :, :k].double().numpy()
    gi, gg = G.grid_pull_backward(dd, xd.requires_grad_(), gd.requires_grad_(), ["dft"], [3], 0)
    ri, rg = O.grid_pull_backward(ds, x0, gs, "dft", 3, 0, oob_f32=True)
    assert _normrel(_np(gg[:1, :k]), rg) <= 1e-5
    dm = torch.zeros_like(dd)
    dm[:1, :, :k] = dd[:1, :, :k]
    gi, gg = G.grid_pull_backward(dm, xd, gd, ["dft"], [3], 0)
    bar = O.grid_push(np.abs(ds), gs, x.shape[2:], "dft", 3, 0, oob_f32=True)
    assert (np.abs(_np(gi[:1]) - ri) <= 1e-4 * bar + 1e-36).all()
    assert float(gi[1:].abs().max()) == 0.0


def _fp64_residual(x, h, g, pen):
    from oracle import solver as S
    xx = x.double().cpu().numpy()
    res = g.double().cpu().numpy() - S.regulariser_grid(xx, h.double().cpu().numpy(), absolute=pen[0],
                                                        membrane=pen[1], bending=pen[2], lame=(pen[3], pen[4]))
    return float(np.linalg.norm(res) / np.linalg.norm(g.double().cpu().numpy()))


@pytest.fixture(scope="module")
def c4():
    from nitorch_amd.bench_data import config4_inputs
    h, g = config4_inputs()
    h = h.to(DEV).movedim(-1, 1).contiguous()
    g = g.to(DEV).movedim(-1, 1).contiguous()
    return h, g, (1e-4, 1e-3, 0.2, 0.05, 0.2)


@pytest.mark.parametrize("use_cg", [False, True])
def test_c4_fmg_full_size(c4, use_cg):
    from nitorch_amd import solve as SV
    h, g, pen = c4
    x, rel, ncyc = SV.solve_to_tolerance(h, g, pen, tol=1e-4, nb_cycles=2, nb_iter=4, max_levels=5, use_cg=use_cg)
    assert rel <= 1e-4
    assert _fp64_residual(x, h, g, pen) <= 1.2e-4


def test_c5_learned_full_size(c4):
    from nitorch_amd import learned as LN
    h, g, pen = c4
    net = LN.UNet.load(LN.default_weights_path(), DEV)
    x, rel, it, fb = LN.pcg_learned(net, h, g, pen, tol=1e-4, max_iter=200)
    assert rel <= 1e-4
    assert _fp64_residual(x, h, g, pen) <= 1.2e-4
