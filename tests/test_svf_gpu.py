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
