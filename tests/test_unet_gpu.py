"""GPU parity of the learned preconditioner's BF16 kernels against the FP64
oracle (oracle/conv_oracle.py), on bf16-rounded inputs and weights.

Tolerances (norm-relative, max|a-b| / max|b|): bf16 outputs 1e-2 (one bf16
rounding of each output, 2^-8 relative, plus FP32 accumulation);
fp32 outputs / FP32-accumulated weight gradients 1e-4; the whole network's
parameter gradients against float64 autograd: cosine >= 0.995 and 1.5e-1
norm-relative (bf16 activations and activation gradients on 11 layers).
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import conv_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel(a, b):
    a = a.detach().float().cpu().double().numpy() if torch.is_tensor(a) else np.asarray(a, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def bf(a):
    """round to bf16, return (device bf16 tensor, fp64 numpy of the rounded values)."""
    t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16)
    return t.to(DEV), t.double().numpy()


@pytest.fixture(scope="module")
def mod():
    from nitorch_amd import learned
    return learned


PAIRS = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (48, 16), (32, 16), (64, 32), (32, 96), (16, 48)]


@pytest.mark.parametrize("ci,co", PAIRS)
def test_conv_fwd_and_weight_grad(mod, ci, co):
    rng = np.random.default_rng(ci * 100 + co)
    net = mod.UNet(DEV)
    for shape in [(3, 5, 20), (4, 4, 16)]:
        x, xr = bf(rng.standard_normal((2,) + shape + (ci,)))
        w = rng.standard_normal((co, 3, 3, 3, ci)) / np.sqrt(27 * ci)
        wk = mod.pack_weight(torch.from_numpy(w).float().to(DEV), ci, co)
        wr = wk[:, :27 * ci].double().cpu().numpy().reshape(co, 3, 3, 3, ci)
        b = rng.standard_normal(co).astype(np.float32)
        bt = torch.from_numpy(b).to(DEV)
        for act in (0, 1):
            y = net._conv(x, shape, ci, co, wk, bt, act)
            ref = O.conv3d_fwd(xr, wr, b, act="lrelu" if act else None)
            assert rel(y, ref) <= 1e-2, (ci, co, shape, act)
        # weight / bias gradients (fp32 accumulation)
        dy, dyr = bf(rng.standard_normal((2,) + shape + (co,)))
        g = {}
        net._bwdw(x, dy, shape, ci, co, g, "t")
        dw, db = g["t"]
        dwr, dbr = O.conv3d_bwd_weight(xr, dyr)
        assert rel(dw[:, :27 * ci].reshape(co, 3, 3, 3, ci), dwr) <= 1e-4, (ci, co, shape)
        assert rel(db, dbr) <= 1e-4, (ci, co, shape)


@pytest.mark.parametrize("ci,co", [(16, 16), (16, 32), (32, 64), (96, 32), (48, 16)])
def test_conv_bwd_data(mod, ci, co):
    """Backward data = forward kernel on the flipped / transposed weights."""
    rng = np.random.default_rng(7 * ci + co)
    net = mod.UNet(DEV)
    shape = (4, 6, 18)
    w = rng.standard_normal((co, 3, 3, 3, ci)) / np.sqrt(27 * ci)
    wt_ = torch.from_numpy(w).float().to(DEV)
    wr = mod.pack_weight(wt_, ci, co)[:, :27 * ci].double().cpu().numpy().reshape(co, 3, 3, 3, ci)
    wt = mod.pack_weight_t(wt_, ci, co)
    dy, dyr = bf(rng.standard_normal((1,) + shape + (co,)))
    dx = net._conv(dy, shape, co, ci, wt, None, 0)
    assert rel(dx, O.conv3d_bwd_data(dyr, wr)) <= 1e-2


def test_head_fp32_and_elementwise(mod):
    rng = np.random.default_rng(3)
    net = mod.UNet(DEV)
    shape = (4, 6, 20)
    x, xr = bf(rng.standard_normal((1,) + shape + (16,)))
    w = rng.standard_normal((3, 3, 3, 3, 16)) * 0.1
    wk = mod.pack_weight(torch.from_numpy(w).float().to(DEV), 16, 16)
    wr = wk[:3, :27 * 16].double().cpu().numpy().reshape(3, 3, 3, 3, 16)
    b = torch.tensor([0.5, -1.0, 2.0] + [0.0] * 13, device=DEV)
    y = net._conv(x, shape, 16, 16, wk, b, 0, out_f32=True, cos=3)
    assert y.shape[-1] == 3 and y.dtype == torch.float32
    assert rel(y, O.conv3d_fwd(xr, wr, b[:3].cpu().numpy())) <= 1e-4
    # pooling, upsampling + concat, leaky-ReLU backward
    a, ar = bf(rng.standard_normal((2, 4, 6, 8, 32)))
    p = net._pool(a, (4, 6, 8), 32)
    assert rel(p, O.avgpool2(ar)) <= 1e-2
    dp = net._pool_bwd(p, (4, 6, 8), 32)
    assert rel(dp, O.avgpool2_bwd(p.double().cpu().numpy())) <= 1e-2
    lo, lor = bf(rng.standard_normal((2, 2, 3, 4, 64)))
    cat = net._upcat(lo, a, (4, 6, 8), 64, 32)
    assert rel(cat, np.concatenate([O.upsample2(lor), ar], -1)) == 0
    dlo, dsk = net._upcat_bwd(cat, (4, 6, 8), 64, 32)
    assert rel(dlo, O.upsample2_bwd(O.upsample2(lor))) <= 1e-2 and rel(dsk, ar) == 0
    dy, dyr = bf(rng.standard_normal(a.shape))
    assert rel(net._lrelu_bwd(dy, a), O.lrelu_bwd(dyr, ar)) <= 1e-2


def _torch_net(params, x):
    """float64 CPU reference of the U-Net (channels-first torch)."""
    def conv(name, t, act=True):
        w = params[name + ".w"].permute(0, 4, 1, 2, 3)
        t = F.conv3d(t, w, params[name + ".b"], padding=1)
        return F.leaky_relu(t, 0.1) if act else t
    a2 = conv("e1b", conv("e1a", x))
    a4 = conv("e2b", conv("e2a", F.avg_pool3d(a2, 2)))
    a6 = conv("b2", conv("b1", F.avg_pool3d(a4, 2)))
    a8 = conv("d2b", conv("d2a", torch.cat([F.interpolate(a6, scale_factor=2), a4], 1)))
    a10 = conv("d1b", conv("d1a", torch.cat([F.interpolate(a8, scale_factor=2), a2], 1)))
    return conv("head", a10, act=False)


def test_network_gradients_vs_float64(mod):
    """Parameter gradients of the whole U-Net (HIP kernels, bf16) against
    float64 autograd on the same bf16-rounded parameters and input."""
    rng = np.random.default_rng(5)
    net = mod.UNet(DEV, seed=3)
    # a non-zero head (it starts at zero), so that every layer gets a gradient
    net.params["head.w"] = torch.randn(net.params["head.w"].shape, generator=torch.Generator().manual_seed(4)
                                       ).to(DEV) * 0.05
    # round the masters to bf16 so both sides use the same weights
    for k in net.params:
        net.params[k] = net.params[k].to(torch.bfloat16).float()
    net.refresh()
    shape = (8, 8, 16)
    g = torch.from_numpy(rng.standard_normal((1, 3) + shape).astype(np.float32)).to(DEV)
    h = torch.from_numpy(rng.uniform(0.5, 2, (1, 6) + shape).astype(np.float32)).to(DEV)
    x = net.pack(g, h)
    out, act = net.forward(x, keep=True)
    target = torch.from_numpy(rng.standard_normal(out.shape).astype(np.float32)).to(DEV)
    dout = 2 * (out - target) / out.numel()
    grads = net.grads_to_params(net.backward(dout, act))
    # float64 reference
    p64 = {k: v.detach().double().cpu().requires_grad_(True) for k, v in net.params.items()}
    x64 = x.double().cpu().permute(0, 4, 1, 2, 3)[:, :9].contiguous()
    out64 = _torch_net(p64, x64)
    assert rel(out.permute(0, 4, 1, 2, 3), out64.detach().numpy()) <= 5e-2
    loss = ((out64 - target.double().cpu().permute(0, 4, 1, 2, 3)) ** 2).mean()
    loss.backward()
    for k in p64:
        a = grads[k].double().cpu().reshape(-1)
        b = p64[k].grad.reshape(-1)
        cos = float(torch.dot(a, b) / (a.norm() * b.norm()).clamp_min(1e-30))
        assert cos >= 0.995, (k, cos)
        assert rel(grads[k], p64[k].grad.numpy()) <= 1.5e-1, k


def test_training_reduces_loss(mod):
    """Adam steps on one small system (forward + backward on the HIP kernels)
    lower the relative error of the network's solution estimate below that
    of the block-Jacobi estimate it corrects."""
    from oracle import solver as S
    rng = np.random.default_rng(9)
    shape = (16, 16, 16)
    g = rng.standard_normal((1, 3) + shape)
    h = rng.uniform(0.5, 2, (1, 6) + shape)
    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
    pen = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2))
    xs, _ = S.cg_grid(h, g, tol=1e-8, **pen)
    net = mod.UNet(DEV, seed=1)
    opt = mod.Adam(net, lr=1e-3)
    gt = torch.from_numpy(g).float().to(DEV)
    ht = torch.from_numpy(h).float().to(DEV)
    xst = torch.from_numpy(xs).float().to(DEV)
    penv = (1e-4, 1e-3, 0.2, 0.05, 0.2)
    losses = [mod.train_step(net, opt, gt, ht, penv, xst)[0] for _ in range(80)]
    base = float((mod.block_jacobi(gt, ht, penv) - xst).pow(2).sum() / xst.pow(2).sum())
    assert abs(losses[0] - base) < 0.05 * base + 1e-3, (losses[0], base)   # head starts near zero
    assert np.mean(losses[-10:]) < 0.7 * base, (base, losses[::10])


def test_pcg_learned_plumbing(mod):
    """The flexible PCG with an untrained network (= block-Jacobi, zero
    head) reaches the tolerance; the residual is checked with the FP64
    oracle operator."""
    from oracle import solver as S
    rng = np.random.default_rng(11)
    shape = (16, 12, 20)
    g = rng.standard_normal((1, 3) + shape)
    h = rng.uniform(0.5, 2, (1, 6) + shape)
    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
    pen = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2))
    net = mod.UNet(DEV, seed=2)
    gt = torch.from_numpy(g).float().to(DEV)
    ht = torch.from_numpy(h).float().to(DEV)
    x, rel_, it, fb = mod.pcg_learned(net, ht, gt, (1e-4, 1e-3, 0.2, 0.05, 0.2), tol=1e-4, max_iter=400)
    assert rel_ <= 1e-4
    res = g - S.regulariser_grid(x.cpu().double().numpy(), h, **pen)
    assert np.linalg.norm(res) / np.linalg.norm(g) <= 2e-4


def test_conv_upcat_matches_upcat_then_conv(mod):
    """ni_unet_conv3_fwd_upcat == ni_unet_up2_concat + ni_unet_conv3_fwd, bit
    for bit, and the network's inference forward (fused) equals its
    training forward (concatenation materialised)."""
    rng = np.random.default_rng(11)
    net = mod.UNet(DEV)
    for shape in [(4, 6, 18), (2, 10, 36), (6, 4, 66)]:
        D, H, W = shape
        a, _ = bf(rng.standard_normal((2, D // 2, H // 2, W // 2, 32)))
        skip, _ = bf(rng.standard_normal((2,) + shape + (16,)))
        w = torch.from_numpy(rng.standard_normal((16, 3, 3, 3, 48)) / np.sqrt(27 * 48)).float().to(DEV)
        wk = mod.pack_weight(w, 48, 16)
        b = torch.from_numpy(rng.standard_normal(16).astype(np.float32)).to(DEV)
        for act in (0, 1):
            ref = net._conv(net._upcat(a, skip, shape, 32, 16), shape, 48, 16, wk, b, act)
            out = net._conv_upcat(a, skip, shape, 32, 16, 16, wk, b, act)
            assert torch.equal(out, ref), (shape, act)
    with torch.no_grad():
        net.params["head.w"].normal_(0, 0.1)
    net.refresh()
    x, _ = bf(rng.standard_normal((1, 8, 8, 16, 16)))
    out_inf = net.forward(x)
    out_train, _ = net.forward(x, keep=True)
    assert out_inf.abs().max() > 0 and torch.equal(out_inf, out_train)
