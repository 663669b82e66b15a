"""GPU parity of scaling and squaring (nitorch_amd/svf.py, repeated HIP
pulls) against the FP64 oracle (oracle/svf.py), and its autodiff gradient
(through the pull's backward kernels) against central differences.
Tolerances: float64 1e-10 relative; float32 1e-5; gradient 1e-6 relative."""
import numpy as np
import pytest
import torch

from oracle import svf as E

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _vel(shape, amp, seed, N=1):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((N,) + shape + (3,))
    F = np.fft.fftn(v, axes=(1, 2, 3))
    k = np.meshgrid(*[np.fft.fftfreq(s) for s in shape], indexing="ij")
    F *= np.exp(-2 * (np.pi * 2) ** 2 * sum(kk ** 2 for kk in k))[None, ..., None]
    v = np.real(np.fft.ifftn(F, axes=(1, 2, 3)))
    return v * amp / np.abs(v).max()


def rel(a, b):
    a = a.detach().cpu().double().numpy()
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("interp,bound,inverse,disp", [("linear", "dft", False, False),
                                                       ("cubic", "dft", True, True),
                                                       ("quadratic", "dct2", False, True)])
def test_exp_vs_oracle(interp, bound, inverse, disp):
    from nitorch_amd import svf
    v = _vel((12, 10, 9), 1.5, 0, N=2)
    ref = E.exp_forward(v, inverse=inverse, steps=6, interpolation=interp, bound=bound, displacement=disp)
    out = svf.exp(torch.from_numpy(v).to(DEV), inverse=inverse, steps=6, interpolation=interp, bound=bound,
                  displacement=disp)
    assert out.shape == v.shape
    assert rel(out, ref) <= 1e-10
    out32 = svf.exp(torch.from_numpy(v).float().to(DEV), inverse=inverse, steps=6, interpolation=interp,
                    bound=bound, displacement=disp)
    assert rel(out32, ref) <= 1e-5


def test_exp_no_batch_and_autograd():
    from nitorch_amd import svf
    v = torch.from_numpy(_vel((10, 9, 8), 1.0, 1)[0]).to(DEV)       # (*spatial, 3)
    w = torch.from_numpy(np.random.default_rng(2).standard_normal(v.shape)).to(DEV)
    vr = v.clone().requires_grad_()
    loss = (svf.exp(vr, steps=5, interpolation="cubic", displacement=True) * w).sum()
    loss.backward()
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(3):
        d = torch.from_numpy(rng.standard_normal(v.shape)).to(DEV)
        f = lambda x: float((svf.exp(x, steps=5, interpolation="cubic", displacement=True) * w).sum())
        fd = (f(v + h * d) - f(v - h * d)) / (2 * h)
        an = float((vr.grad * d).sum())
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-9, (fd, an)


# --- Jacobians, analytic backward (exp_backward, anagrad=True) --------------
@pytest.mark.parametrize("bound", ["dft", "dct2", "dct1", "dst2", "zero", "replicate"])
def test_grid_jacobian_vs_oracle(bound):
    from nitorch_amd import spatial as SP
    v = _vel((9, 7, 8), 1.2, 10, N=2)
    ref = E.grid_jacobian(v, bound=bound, type="disp")
    out = SP.grid_jacobian(torch.from_numpy(v).to(DEV), bound=bound, type="disp")
    assert out.shape == ref.shape and rel(out, ref) <= 1e-12
    ident = E.identity((9, 7, 8))[None]
    out = SP.grid_jacobian(torch.from_numpy(v + ident).to(DEV), bound=bound)
    assert rel(out, ref) <= 1e-12


def test_grid_jacobian_sampled():
    """With `sample`, the Jacobian is the spline derivative (grid_grad) of the
    displacement at the sample points, plus the identity."""
    from nitorch_amd import spatial as SP
    from oracle import pushpull as PP
    v = _vel((8, 9, 7), 1.0, 11)
    s = E.identity((8, 9, 7))[None] + _vel((8, 9, 7), 0.8, 12)
    ref = PP.grid_grad(np.ascontiguousarray(np.moveaxis(v, -1, 1)), np.ascontiguousarray(s), "dft", 1, 1)
    ref = np.moveaxis(ref, 1, -2) + np.eye(3)
    out = SP.grid_jacobian(torch.from_numpy(v).to(DEV), sample=torch.from_numpy(s).to(DEV), type="disp",
                           bound="dft", extrapolate=True)
    assert rel(out, ref) <= 1e-10


def test_exp_forward_jacobian_vs_oracle():
    from nitorch_amd import svf
    v = _vel((12, 10, 9), 1.5, 13, N=2)
    dref, jref = E.exp_forward_jac(v, steps=4)
    d, j = svf.exp_forward(torch.from_numpy(v).to(DEV), steps=4, displacement=True, jacobian=True)
    assert rel(d, dref) <= 1e-10 and rel(j, jref) <= 1e-10


@pytest.mark.parametrize("inverse,rotate,with_hess", [(False, False, False), (True, False, True),
                                                      (False, True, True), (True, True, False)])
def test_exp_backward_vs_oracle(inverse, rotate, with_hess):
    from nitorch_amd import svf
    rng = np.random.default_rng(14)
    v = _vel((11, 9, 10), 1.3, 15)
    g = rng.standard_normal(v.shape)
    h = rng.standard_normal(v.shape[:-1] + (6,)) if with_hess else None
    ref = E.exp_backward(v, g, h, inverse=inverse, steps=4, rotate_grad=rotate) if with_hess else \
        E.exp_backward(v, g, inverse=inverse, steps=4, rotate_grad=rotate)
    args = [torch.from_numpy(g).to(DEV)] + ([torch.from_numpy(h).to(DEV)] if with_hess else [])
    out = svf.exp_backward(torch.from_numpy(v).to(DEV), *args, inverse=inverse, steps=4, rotate_grad=rotate)
    if with_hess:
        assert rel(out[0], ref[0]) <= 1e-9 and rel(out[1], ref[1]) <= 1e-9
    else:
        assert rel(out, ref) <= 1e-9


def test_anagrad_autograd():
    """exp(anagrad=True): forward equals exp_forward, backward equals
    exp_backward(rotate_grad=True) of the upstream gradient."""
    from nitorch_amd import svf
    v = torch.from_numpy(_vel((10, 9, 8), 1.1, 16)).to(DEV).requires_grad_(True)
    g = torch.randn(v.shape, dtype=v.dtype, device=DEV)
    y = svf.exp(v, steps=5, anagrad=True)
    assert rel(y, svf.exp_forward(v.detach(), steps=5).cpu().numpy()) <= 1e-12
    (gv,) = torch.autograd.grad(y, v, g)
    ref = E.exp_backward(v.detach().cpu().numpy(), g.cpu().numpy(), steps=5, rotate_grad=True)
    assert rel(gv, ref) <= 1e-9


def test_exp_backward_full_size_closed_form():
    """160 x 192 x 160, float32: a constant velocity of integer steps u per
    squaring gives the average of the gradient rolled by m u (dft, linear)."""
    from nitorch_amd import svf
    s, u = 3, (1, 2, -1)
    v = torch.tensor([c * 2.0 ** s for c in u], device=DEV).expand(1, 160, 192, 160, 3).contiguous()
    g = torch.randn(v.shape, device=DEV)
    out = svf.exp_backward(v, g, steps=s, rotate_grad=True)
    ref = sum(torch.roll(g, shifts=tuple(m * c for c in u), dims=(1, 2, 3)) for m in range(2 ** s)) / 2 ** s
    assert (out - ref).abs().max().item() <= 1e-5 * g.abs().max().item()


@pytest.mark.parametrize("dt,tol", [(torch.float64, 1e-12), (torch.float32, 2e-5)])
def test_mat3_kernels_vs_torch(dt, tol):
    """ni_svf_mat3_mul / ni_svf_grad_step against float64 torch products, on
    dense, transposed, channel-first (entries as channels) and stride-0
    (expanded identity) views."""
    from nitorch_amd import svf
    g = torch.Generator(device="cpu").manual_seed(5)
    shp = (2, 5, 6, 7)
    a = torch.randn(shp + (3, 3), generator=g, dtype=torch.float64)
    b = torch.randn(shp + (3, 3), generator=g, dtype=torch.float64)
    p = torch.randn(shp + (3,), generator=g, dtype=torch.float64)
    g0 = torch.randn(shp + (3,), generator=g, dtype=torch.float64)
    A, B, P, G0 = (t.to(DEV, dt) for t in (a, b, p, g0))

    def rel(x, ref):
        return float((x.double().cpu() - ref).abs().max() / ref.abs().max())

    def chan_first(t, k):    # same values, entries stored as leading channels
        return t.reshape(shp + (k,)).movedim(-1, 1).contiguous().movedim(1, -1)

    assert rel(svf._matmul(A, B), a @ b) <= tol
    assert rel(svf._matmul(A.transpose(-1, -2), B), a.transpose(-1, -2) @ b) <= tol
    Bc = chan_first(B, 9).reshape(shp + (3, 3))
    assert Bc.stride(-1) != 1 and rel(svf._matmul(A, Bc), a @ b) <= tol
    eye = torch.eye(3, dtype=dt, device=DEV).expand(shp + (3, 3))
    assert torch.equal(svf._matmul(A, eye), A)
    ref = g0 + torch.linalg.det(a)[..., None] * (a.transpose(-1, -2) @ p[..., None])[..., 0]
    out = svf._grad_step(A, chan_first(P, 3), G0)
    assert out is not None and rel(out, ref) <= tol
    # not applicable: mismatched shapes fall back to the caller's torch path
    assert svf._grad_step(A, P[:1], G0) is None
