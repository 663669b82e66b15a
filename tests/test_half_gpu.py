"""Half-precision inputs on the GPU path (the reference's GPU backend
dispatches float16, nitorch/_C/src/impl/pushpull_common.cpp:2505-2537):
float16 tensors run with float32 accumulation and return float16.  Checked
against the FP64 oracle on the float16-rounded inputs; tolerance 1e-3
norm-relative (the outputs are rounded to float16, 2^-11 relative)."""
import numpy as np
import pytest
import torch

from oracle import pushpull as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = a.detach().cpu().double().numpy()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("order", [1, 3])
def test_half_pull_push_count_grad(order):
    from nitorch_amd import grid as G
    rng = np.random.default_rng(1200 + order)
    shape = (14, 12, 17)
    ident = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij"), -1)[None]
    g = (ident + rng.standard_normal((2,) + shape + (3,))).astype(np.float16)
    x = rng.standard_normal((2, 3) + shape).astype(np.float16)
    y = rng.standard_normal((2, 3) + shape).astype(np.float16)
    g64, x64, y64 = g.astype(np.float64), x.astype(np.float64), y.astype(np.float64)
    gt, xt, yt = (torch.from_numpy(a).to(DEV) for a in (g, x, y))
    for bound in ("dct2", "dft"):
        out = G.grid_pull(xt, gt, [bound], [order], 0)
        assert out.dtype == torch.float16
        assert rel(out, O.grid_pull(x64, g64, bound, order, 0)) <= 1e-3
        out = G.grid_push(yt, gt, list(shape), [bound], [order], 0)
        assert out.dtype == torch.float16
        assert rel(out, O.grid_push(y64, g64, shape, bound, order, 0)) <= 1e-3
        out = G.grid_count(gt, list(shape), [bound], [order], 0)
        assert rel(out, O.grid_count(g64, shape, bound, order, 0)) <= 1e-3
        out = G.grid_grad(xt, gt, [bound], [order], 0)
        assert rel(out, O.grid_grad(x64, g64, bound, order, 0)) <= 1e-3
        # backward: both gradients, float16 in and out
        xr, gr = xt.clone().requires_grad_(), gt.clone().requires_grad_()
        gi, gg = G.grid_pull_backward(yt, xr, gr, [bound], [order], 0)
        ri, rg = O.grid_pull_backward(y64, x64, g64, bound, order, 0)
        assert gi.dtype == torch.float16 and gg.dtype == torch.float16
        assert rel(gi, ri) <= 1e-3 and rel(gg, rg) <= 2e-3
