"""GPU parity of the spline prefilter (HIP `ni_spline_coeff`, through
nitorch_amd.spatial.spline_coeff / spline_coeff_nd) against the FP64 oracle
(oracle/coeff.py): every order 0-7 with the zero, dct1, replicate, dct2 and
dft bounds, lines of length 1, 2 and odd lengths, lines long enough for the
global-memory path (float > 630, double > 315 samples), non-contiguous and
in-place calls.  At the 4 x 8 x 192^3 pull/push size the size-independent
check is the interpolation property: pull(prefilter=True) at the identity
grid returns the image.
Tolerances (max abs error / max abs value): float64 1e-11, float32 2e-6
(LDS path, one rounding per recursion step) and 1e-5 at 192^3."""
import numpy as np
import pytest
import torch

from oracle import coeff as C

pytestmark = pytest.mark.gpu
DEV = "cuda"
BOUNDS = ["zero", "dct1", "replicate", "dct2", "dft"]


def rel(a, b):
    a = a.detach().cpu().double().numpy()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def tol(dt):
    return 1e-11 if dt == torch.float64 else 2e-6


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
@pytest.mark.parametrize("order", list(range(8)))
def test_nd_vs_oracle(dt, order):
    from nitorch_amd import spatial as SP
    rng = np.random.default_rng(order)
    for shape in [(5, 1, 7), (2, 3, 4), (17, 12, 33), (64, 9, 70)]:
        x = rng.standard_normal((2, 3) + shape)
        for bound in BOUNDS:
            ref = C.spline_coeff_nd(x, bound, order, dim=3)
            out = SP.spline_coeff_nd(torch.from_numpy(x).to(DEV, dt), interpolation=order, bound=bound, dim=3)
            assert out.shape == x.shape and out.dtype == dt
            assert rel(out, ref) <= tol(dt), (shape, bound)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_mixed_axes_and_long_lines(dt):
    """Per-axis bounds and orders; lines too long for LDS take the global
    path (float: 700 > 630; double: 330 > 315)."""
    from nitorch_amd import spatial as SP
    rng = np.random.default_rng(11)
    n = 700 if dt == torch.float32 else 330
    for shape, bound, order in [((n, 3, 2), ["dct2", "dft", "dct1"], [3, 5, 7]),
                                ((3, n, 2), ["dft"], [2, 4]),
                                ((2, 3, n), ["zero", "replicate"], [7, 1, 6])]:
        x = rng.standard_normal((2,) + shape)
        ref = C.spline_coeff_nd(x, bound, order, dim=3)
        out = SP.spline_coeff_nd(torch.from_numpy(x).to(DEV, dt), interpolation=order, bound=bound, dim=3)
        assert rel(out, ref) <= tol(dt) * (3 if dt == torch.float32 else 1), (shape, bound, order)


def test_one_dim_noncontiguous_and_inplace():
    from nitorch_amd import spatial as SP
    rng = np.random.default_rng(12)
    x = rng.standard_normal((3, 10, 11, 12))
    for dim in range(-4, 4):
        for bound in ("dct2", "dft"):
            ref = C.filter1d(x, bound, 3, axis=dim)
            t = torch.from_numpy(x).to(DEV)
            out = SP.spline_coeff(t, interpolation=3, bound=bound, dim=dim)
            assert rel(out, ref) <= 1e-11 and np.array_equal(t.cpu().numpy(), x)
    # non-contiguous input, in place: the caller's tensor holds the result
    base = torch.from_numpy(x.transpose(0, 2, 1, 3).copy()).to(DEV)
    view = base.transpose(1, 2)                      # shape of x, not contiguous
    ret = SP.spline_coeff_nd(view, interpolation=5, bound="dct1", dim=3, inplace=True)
    ref = C.spline_coeff_nd(x, "dct1", 5, dim=3)
    assert ret.data_ptr() == view.data_ptr() and rel(view, ref) <= 1e-11
    # contiguous in place: same storage
    t = torch.from_numpy(x).to(DEV)
    ret = SP.spline_coeff_nd(t, interpolation=2, bound="dft", dim=2, inplace=True)
    assert ret.data_ptr() == t.data_ptr() and rel(t, C.spline_coeff_nd(x, "dft", 2, dim=2)) <= 1e-11


def test_errors_and_identities():
    from nitorch_amd import spatial as SP
    t = torch.randn(2, 5, 6, 7, device=DEV)
    with pytest.raises(NotImplementedError):
        SP.spline_coeff_nd(t, interpolation=3, bound="dst2", dim=3)
    # orders 0/1 ignore the bound, as the reference returns before using it
    assert torch.equal(SP.spline_coeff_nd(t, interpolation=1, bound="dst2", dim=3), t)
    assert SP.spline_coeff_nd(torch.zeros(0, 4, 4, 4, device=DEV), interpolation=3).shape == (0, 4, 4, 4)
    with pytest.raises(RuntimeError):
        SP.spline_coeff_nd(torch.randn(2, 5, 6, 7), interpolation=3)   # CPU tensor: no CPU path


def test_backward_is_forward_of_grad():
    from nitorch_amd import spatial as SP
    x = torch.randn(2, 9, 8, 7, device=DEV, dtype=torch.float64, requires_grad=True)
    g = torch.randn(2, 9, 8, 7, device=DEV, dtype=torch.float64)
    y = SP.spline_coeff_nd(x, interpolation=3, bound="dct2", dim=3)
    (gx,) = torch.autograd.grad(y, x, g)
    ref = C.spline_coeff_nd(g.cpu().numpy(), "dct2", 3, dim=3)
    assert rel(gx, ref) <= 1e-11
    y1 = SP.spline_coeff(x, interpolation=4, bound="dft", dim=1)
    (gx,) = torch.autograd.grad(y1, x, g)
    assert rel(gx, C.filter1d(g.cpu().numpy(), "dft", 4, axis=1)) <= 1e-11


@pytest.mark.parametrize("bound,order", [("dct2", 3), ("dft", 5), ("dct1", 7)])
def test_prefiltered_pull_interpolates(bound, order):
    from nitorch_amd import spatial as SP
    torch.manual_seed(order)
    x = torch.randn(2, 3, 40, 37, 33, device=DEV, dtype=torch.float64)
    grid = SP.identity_grid(x.shape[2:], dtype=x.dtype, device=DEV)
    y = SP.grid_pull(x, grid, interpolation=order, bound=bound, prefilter=True)
    assert (y - x).abs().max().item() <= 1e-9 * x.abs().max().item()
    # push(prefilter=True) = prefilter of the push
    z = SP.grid_push(x, grid, interpolation=order, bound=bound, prefilter=True)
    ref = C.spline_coeff_nd(SP.grid_push(x, grid, interpolation=order, bound=bound).cpu().numpy(),
                            bound, order, dim=3)
    assert rel(z, ref) <= 1e-11


def test_prefilter_full_size():
    """4 x 8 x 192^3 float32, cubic, dct2: the prefiltered pull at the
    identity grid returns the image (the workload of the pull/push metric)."""
    from nitorch_amd import spatial as SP
    g = torch.Generator(device="cpu").manual_seed(2)
    x = torch.rand(4, 8, 192, 192, 192, generator=g).to(DEV)
    grid = SP.identity_grid(x.shape[2:], dtype=x.dtype, device=DEV)
    c = SP.spline_coeff_nd(x, interpolation=3, bound="dct2", dim=3)
    y = SP.grid_pull(c, grid, interpolation=3, bound="dct2")
    err = (y - x).abs().max().item()
    assert err <= 1e-5, err
    # one line of the volume against the oracle
    ref = C.spline_coeff_nd(x[1, 5].double().cpu().numpy(), "dct2", 3, dim=3)
    assert rel(c[1, 5], ref) <= 1e-5
