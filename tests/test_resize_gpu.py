"""GPU parity of the compiled resize (`ni_resize` through
nitorch_amd.resize) against the FP64 oracle (oracle/resize.py, itself
pinned against the reference's compiled resize), and of the spatial-level
resize / resize_grid / reslice (prefiltered grid_pull on affine lattices)
against the oracle's prefilter + pull.  Tolerances (max abs error / max
abs value): float64 1e-10, float32 2e-5.  At the 4 x 8 x 192^3 size the
checks are size-independent: the adjoint identity <R x, y> = <x, R^T y>
and agreement of the resize-based prolongation with the FMG one."""
import numpy as np
import pytest
import torch

from oracle import coeff as CF
from oracle import pushpull as PP
from oracle import resize as RS

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = a.detach().cpu().double().numpy() if torch.is_tensor(a) else a
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def T(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dt)


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("order", [[0], [1], [2], [3], [5], [7], [1, 3, 0]])
def test_resize_vs_oracle(dt, order):
    from nitorch_amd import resize as RZ
    tol = 1e-10 if dt == torch.float64 else 2e-5
    rng = np.random.default_rng(len(order) * 10 + order[0])
    for mode in (["edge"], ["center"], ["first"], ["last"], ["e", "c", "f"]):
        for f in ([2.0], [0.5], [1.5], [0.7, 1.3, 2.0]):
            for adjoint in (False, True):
                for normalize in (False, True):
                    x = rng.standard_normal((2, 3, 7, 6, 5))
                    bound = ["dct2", "dft", "replicate"]
                    ref = RS.resize(x, f, bound, order, [RS.mode_code(m) for m in mode], None, None, adjoint,
                                    normalize)
                    out = RZ.resize(T(x, dt), None, f, bound, order, mode, adjoint, normalize)
                    assert out.shape == ref.shape and out.dtype == dt
                    assert rel(out, ref) <= tol, (mode, f, adjoint, normalize)


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_resize_long_rows_vs_oracle(dt):
    """Rows of 64 or more values per pass (the row-per-work-group kernel)."""
    from nitorch_amd import resize as RZ
    tol = 1e-10 if dt == torch.float64 else 2e-5
    rng = np.random.default_rng(21)
    x = rng.standard_normal((1, 2, 12, 40, 70))
    for order, mode, f in (([3], ["edge"], [1.5]), ([0], ["center"], [0.6]), ([5, 1, 2], ["first"], [0.8])):
        for given in (False, True):
            shp = RS.out_shape(x.shape[2:], f, False)
            o = rng.standard_normal((1, 2) + tuple(shp)) if given else None
            ref = RS.resize(x, f, ["dft", "dct2", "replicate"], order, [RS.mode_code(m) for m in mode], None, o)
            out = RZ.resize(T(x, dt), None if o is None else T(o, dt), f, ["dft", "dct2", "replicate"], order,
                            mode)
            assert rel(out, ref) <= tol, (order, mode, f, given)


def test_given_output_and_other_dims():
    from nitorch_amd import resize as RZ
    rng = np.random.default_rng(7)
    # a given output: added into by the generic kernel, overwritten by the
    # isotropic linear one, zeroed for the adjoint
    x = rng.standard_normal((1, 2, 6, 7, 5))
    for order, adjoint in (([3], False), ([1], False), ([1, 2, 3], False), ([3], True)):
        shp = RS.out_shape(x.shape[2:], [1.5], adjoint)
        o = rng.standard_normal((1, 2) + tuple(shp))
        ref = RS.resize(x, [1.5], "dft", order, [0], None, o, adjoint, False)
        out = T(o)
        ret = RZ.resize(T(x), out, [1.5], ["dft"], order, ["edge"], adjoint, False)
        assert ret.data_ptr() == out.data_ptr() and rel(out, ref) <= 1e-10, (order, adjoint)
    # 1-D and 2-D lattices
    for shp in [(9,), (8, 5)]:
        x = rng.standard_normal((2, 3) + shp)
        for adjoint in (False, True):
            ref = RS.resize(x, [1.7], "dct2", [3], [0], None, None, adjoint, True)
            out = RZ.resize(T(x), None, [1.7], ["dct2"], [3], ["edge"], adjoint, True)
            assert rel(out, ref) <= 1e-10, (shp, adjoint)


def test_c_wrappers_and_backward():
    from nitorch_amd import resize as RZ
    from nitorch_amd import solve as SV
    rng = np.random.default_rng(8)
    x = T(rng.standard_normal((2, 8, 9, 8, 10)))     # more channels than the FMG's 6
    # c_resize = resize with the reference's argument handling
    ref = RS.resize(x.cpu().numpy(), [1.5], "dft", [2], [RS.FIRST])
    assert rel(RZ.c_resize(x, 1.5, "dft", 2, "first"), ref) <= 1e-10
    ref = RS.resize(x.cpu().numpy(), None, "dct2", [1], [RS.CENTER], shape=(5, 6, 7))
    assert rel(RZ.c_resize(x, shape=(5, 6, 7)), ref) <= 1e-10
    # c_prolongation / c_restriction are the FMG transfers
    assert torch.equal(RZ.c_prolongation(x), SV.fmg_prolongation(x, torch.zeros(2, 8, 18, 16, 20, device=DEV,
                                                                             dtype=x.dtype)))
    assert torch.equal(RZ.c_restriction(x), SV.fmg_restriction(x, torch.zeros(2, 8, 5, 4, 5, device=DEV,
                                                                            dtype=x.dtype)))
    # the resize-based prolongation / restriction agree with them
    assert rel(RZ.prolongation(x, torch.zeros(2, 8, 18, 16, 20, device=DEV, dtype=x.dtype)),
               RZ.c_prolongation(x).cpu().numpy()) <= 1e-10
    assert rel(RZ.restriction(x, torch.zeros(2, 8, 5, 4, 5, device=DEV, dtype=x.dtype)),
               RZ.c_restriction(x).cpu().numpy()) <= 1e-10
    # backward = transpose
    for order in ([1], [3], [0, 2, 4]):
        y = RZ.resize(x, None, [1.6], ["dct2"], order, ["first"])
        g = torch.randn_like(y)
        back = RZ.resize_backward(g, torch.zeros_like(x), [1.6], ["dct2"], order, ["first"])
        lhs, rhs = (y * g).sum().item(), (x * back).sum().item()
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # grid variants rescale the components as the reference does
    d = T(rng.standard_normal((1, 6, 5, 4, 3)))
    p = RZ.c_prolongation_grid(d)
    base = RZ.c_prolongation(d.movedim(-1, 1)).movedim(1, -1)
    fac = [o / i for i, o in zip((6, 5, 4), p.shape[-3:])]
    for c in range(3):
        assert torch.allclose(p[..., c], base[..., c] * fac[c])


def _oracle_resize_spatial(x, outshape, order, bound, lin):
    c = CF.spline_coeff_nd(x, bound, order, dim=3)
    grid = np.stack(np.meshgrid(*lin, indexing="ij"), -1)[None]
    grid = np.ascontiguousarray(np.broadcast_to(grid, (x.shape[0],) + grid.shape[1:]))
    return PP.grid_pull(c, grid, bound, order, 0)


def test_spatial_resize_resize_grid_reslice():
    from nitorch_amd import spatial as SP
    from nitorch_amd.affine import affine_resize
    rng = np.random.default_rng(10)
    x = rng.standard_normal((2, 3, 12, 10, 9))
    # anchors: centers / edges / first / last
    for anchor in ("c", "e", "f", "l"):
        out = SP.resize(T(x), factor=[1.5, 0.8, 2.0], anchor=anchor, interpolation=3, bound="dct2")
        outshape = [int(s * f) for s, f in zip(x.shape[2:], [1.5, 0.8, 2.0])]
        lin = []
        for n, m, f in zip(x.shape[2:], outshape, [1.5, 0.8, 2.0]):
            if anchor == "c":
                lin.append(np.linspace(0, n - 1, m))
            elif anchor == "e":
                lin.append(np.arange(m) * (n / m) + 0.5 * (n / m - 1))
            elif anchor == "f":
                lin.append(np.arange(m) / f)
            else:
                lin.append(np.arange(m) / f + (n - 1) - (m - 1) / f)
        ref = _oracle_resize_spatial(x, outshape, 3, "dct2", lin)
        assert out.shape[2:] == tuple(outshape) and rel(out, ref) <= 1e-9, anchor
    # factor 1 with prefiltering returns the image (dct1: exact interpolation
    # on lines this short, where the reference's dct2 start value is off by
    # O(z^(2n-1)))
    same = SP.resize(T(x), factor=1, interpolation=5, bound="dct1")
    assert rel(same, x) <= 1e-9
    # affine output composes the voxel scaling
    aff = torch.eye(4, dtype=torch.float64)
    aff[:3, :3] *= 2.0
    _, aff2 = SP.resize(T(x), factor=2, affine=aff.to(DEV), interpolation=1)
    ref_aff, _ = affine_resize(aff, x.shape[2:], 2, "c")
    assert torch.allclose(aff2.cpu(), ref_aff)
    # resize_grid: a constant displacement is rescaled by the size ratio
    disp = torch.full((1, 12, 10, 9, 3), 1.5, dtype=torch.float64, device=DEV)
    r = SP.resize_grid(disp, shape=(24, 20, 18), type="displacement", anchor="e", interpolation=1,
                       bound="dct2", extrapolate=True)
    assert torch.allclose(r, torch.full_like(r, 3.0))
    # reslice: identity is the image; an integer translation with dft is a roll
    a = torch.eye(4, dtype=torch.float64, device=DEV)
    same = SP.reslice(T(x), a, a, interpolation=3, bound="dft", extrapolate=True)
    assert rel(same, x) <= 1e-9
    b = a.clone()
    b[0, 3] = 2.0
    moved = SP.reslice(T(x), a, b, interpolation=3, bound="dft", extrapolate=True)
    assert rel(moved, np.roll(x, -2, axis=2)) <= 1e-9


def test_full_size_adjoint_and_prolongation():
    """4 x 8 x 192^3 float32: resize by 0.5 (cubic, dct2) and its adjoint
    are transposes; the quadratic edge-aligned resize equals the FMG
    prolongation at 96^3 -> 192^3."""
    from nitorch_amd import resize as RZ
    g = torch.Generator(device="cpu").manual_seed(3)
    x = torch.rand(4, 8, 192, 192, 192, generator=g).to(DEV)
    y = RZ.resize(x, None, [0.5], ["dct2"], [3], ["edge"])
    assert y.shape == (4, 8, 96, 96, 96)
    w = torch.rand(y.shape, generator=g).to(DEV)
    back = RZ.resize_backward(w, torch.zeros_like(x), [0.5], ["dct2"], [3], ["edge"])
    lhs = (y.double() * w.double()).sum().item()
    rhs = (x.double() * back.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    p1 = RZ.prolongation(y, torch.zeros_like(x))
    p2 = RZ.c_prolongation(y)
    assert rel(p1, p2.double().cpu().numpy()) <= 1e-5


def test_resize_view_at_odd_offset():
    """A contiguous input that starts one element into its storage (not
    aligned to 16 bytes) gives the same result as an aligned copy."""
    from nitorch_amd import resize as RZ
    shape = (1, 2, 12, 16, 72)
    n = int(np.prod(shape))
    big = torch.randn(n + 1, device=DEV)
    x = big[1:].view(shape)
    assert x.is_contiguous() and x.data_ptr() % 16 != 0
    a = RZ.resize(x, None, [0.5], ["dct2"], [3], ["edge"])
    b = RZ.resize(x.clone(), None, [0.5], ["dct2"], [3], ["edge"])
    assert torch.equal(a, b)
