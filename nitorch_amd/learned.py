"""Learned preconditioner for the displacement-field Gauss-Newton system
(H + L) x = g: a small 3-D U-Net that maps (g, H) to an approximation of
x, used as warm start and as the preconditioner of a flexible conjugate
gradient, with full multigrid as the fallback when the residual stalls.

No counterpart in the reference (nitorch solves the system with FMG /
PCG only, nitorch/_C/src/impl/fmg_common.cpp); the operator L, the residual
and the fallback are the reference-equivalent kernels of `nitorch_amd.solve`.

Network (3 levels, 16 / 32 / 64 channels, 3x3x3 convolutions, leaky ReLU,
BF16 weights and activations, FP32 accumulation):

    in(16: g, H, 0-padding) -> e1a 16 -> e1b 16 ---------------------------> concat -> d1a 16 -> d1b 16 -> head 3
                                  \\ pool -> e2a 32 -> e2b 32 -----> concat -> d2a 32 -> d2b 32 -/ up
                                                     \\ pool -> b1 64 -> b2 64 -/ up

The network predicts a correction to the block-Jacobi estimate
(H + diag L)^-1 g (the reference's precond_grid), on the input normalised by
rms(g); its last layer starts at zero, so the untrained preconditioner is
block-Jacobi.  Every convolution, its backward passes, pooling,
upsampling+concat and the input / output packing run on the HIP kernels of
libnitorch_amd.so (`ni_unet_*`); torch holds the buffers and runs the
optimiser update.  Spatial sizes must be multiples of 4.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from . import solve as SV

_F = ctypes.c_float
_PF = ctypes.POINTER(ctypes.c_float)

L.register_signatures({
    "ni_unet_conv3_fwd": [L._I64, L._PI64, L._I32, L._I32, L._P, L._P, L._P, L._P, L._I32, L._I32, L._I32, L._P],
    "ni_unet_conv3_fwd_upcat": [L._I64, L._PI64, L._I32, L._I32, L._I32, L._P, L._P, L._P, L._P, L._P, L._I32, L._P],
    "ni_unet_conv3_bwd_weight": [L._I64, L._PI64, L._I32, L._I32, L._P, L._P, L._P, L._P, L._P],
    "ni_unet_lrelu_bwd": [L._I64, L._P, L._P, L._P, L._P],
    "ni_unet_avgpool2": [L._I32, L._I64, L._PI64, L._I32, L._P, L._P, L._P],
    "ni_unet_up2_concat": [L._I32, L._I64, L._PI64, L._I32, L._I32, L._P, L._P, L._P, L._P],
    "ni_unet_pack_input": [L._I64, L._I64, L._P, L._P, L._I32, L._P, _F, _F, L._P],
    "ni_unet_unpack_output": [L._I64, L._I64, L._P, L._P, _F, L._I32, L._P],
})

# (name, CI real, CI stored, CO real, CO stored, activation)
LAYERS = [
    ("e1a", 9, 16, 16, 16, 1), ("e1b", 16, 16, 16, 16, 1),
    ("e2a", 16, 16, 32, 32, 1), ("e2b", 32, 32, 32, 32, 1),
    ("b1", 32, 32, 64, 64, 1), ("b2", 64, 64, 64, 64, 1),
    ("d2a", 96, 96, 32, 32, 1), ("d2b", 32, 32, 32, 32, 1),
    ("d1a", 48, 48, 16, 16, 1), ("d1b", 16, 16, 16, 16, 1),
    ("head", 16, 16, 3, 16, 0),
]


def _kp(ci):
    return (27 * ci + 31) // 32 * 32


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _shape(d, h, w):
    return L.i64arr((d, h, w))


def _check(rc, what):
    L.check(rc, what)


def pack_weight(w, ci_s, co_s):
    """fp32 [CO, 3, 3, 3, CI] -> bf16 [co_s, KP] with K = tap * ci_s + ci."""
    co, ci = w.shape[0], w.shape[-1]
    wp = torch.zeros(co_s, 27, ci_s, dtype=torch.float32, device=w.device)
    wp[:co, :, :ci] = w.reshape(co, 27, ci)
    out = torch.zeros(co_s, _kp(ci_s), dtype=torch.bfloat16, device=w.device)
    out[:, :27 * ci_s] = wp.reshape(co_s, 27 * ci_s).to(torch.bfloat16)
    return out


def pack_weight_t(w, ci_s, co_s):
    """Weights of the backward-data convolution: flipped taps, channels
    swapped: bf16 [ci_s, KP'] with K' = tap' * co_s + co, tap' = 26 - tap."""
    co, ci = w.shape[0], w.shape[-1]
    wp = torch.zeros(co_s, 27, ci_s, dtype=torch.float32, device=w.device)
    wp[:co, :, :ci] = w.reshape(co, 27, ci)
    wt = wp.flip(1).permute(2, 1, 0).reshape(ci_s, 27 * co_s)       # [ci_s, tap', co_s]
    out = torch.zeros(ci_s, _kp(co_s), dtype=torch.bfloat16, device=w.device)
    out[:, :27 * co_s] = wt.to(torch.bfloat16)
    return out


class UNet:
    """The U-Net's parameters (fp32 masters) and its forward / backward on
    the HIP kernels.  Inputs are channel-first fp32 g [N, 3, D, H, W] and
    H [N, CC, D, H, W] (the solver layout); the output is fp32 [N, 3, D, H, W]."""

    def __init__(self, device="cuda", seed=0):
        gen = torch.Generator().manual_seed(seed)
        self.device = torch.device(device)
        self.params = {}
        for name, ci, ci_s, co, co_s, act in LAYERS:
            fan_in = 27 * ci
            gain = math.sqrt(2.0 / (1 + 0.1 ** 2)) if act else 1.0
            std = gain / math.sqrt(fan_in)
            w = torch.randn(co, 3, 3, 3, ci, generator=gen) * std
            if name == "head":   # the network starts as the identity correction (zero)
                w.zero_()
            self.params[name + ".w"] = w.to(self.device)
            self.params[name + ".b"] = torch.zeros(co, device=self.device)
        self.refresh()

    # -- weights --------------------------------------------------------------
    def refresh(self):
        """Re-pack the bf16 kernel weights from the fp32 masters."""
        self.wk, self.wt, self.bias = {}, {}, {}
        for name, ci, ci_s, co, co_s, act in LAYERS:
            w = self.params[name + ".w"]
            self.wk[name] = pack_weight(w, ci_s, co_s)
            self.wt[name] = pack_weight_t(w, ci_s, co_s)
            b = torch.zeros(co_s, device=self.device)
            b[:co] = self.params[name + ".b"]
            self.bias[name] = b

    def state_dict(self):
        return {k: v.detach().cpu().numpy() for k, v in self.params.items()}

    def load_state_dict(self, sd):
        for k in self.params:
            self.params[k] = torch.as_tensor(np.asarray(sd[k]), dtype=torch.float32, device=self.device)
        self.refresh()

    def save(self, path):
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path, device="cuda"):
        net = cls(device)
        with np.load(path, allow_pickle=False) as f:
            net.load_state_dict({k: f[k] for k in f.files})
        return net

    # -- primitives -----------------------------------------------------------
    def _conv(self, x, shape, ci, co, w, bias, act, out_f32=False, cos=None):
        N, (D, H, W) = x.shape[0], shape
        cos = co if cos is None else cos
        if out_f32:
            y = torch.empty(N, D, H, W, cos, dtype=torch.float32, device=x.device)
        else:
            y = torch.empty(N, D, H, W, co, dtype=torch.bfloat16, device=x.device)
        _check(L.lib().ni_unet_conv3_fwd(N, _shape(D, H, W), ci, co, _ptr(x), _ptr(w), _ptr(bias), _ptr(y),
                                         int(act), int(out_f32), int(cos), L.stream_of(x)), "unet_conv3_fwd")
        return y

    def _conv_upcat(self, a, skip, shape, c1, c2, co, w, bias, act):
        """conv(concat(up2(a), skip)) in one pass (bf16 out)."""
        N = a.shape[0]
        D, H, W = shape
        y = torch.empty(N, D, H, W, co, dtype=torch.bfloat16, device=a.device)
        _check(L.lib().ni_unet_conv3_fwd_upcat(N, _shape(D, H, W), c1, c2, co, _ptr(a), _ptr(skip), _ptr(w),
                                               _ptr(bias), _ptr(y), int(act), L.stream_of(a)),
               "unet_conv3_fwd_upcat")
        return y

    def _pool(self, x, shape, C):
        N, (D, H, W) = x.shape[0], shape
        y = torch.empty(N, D // 2, H // 2, W // 2, C, dtype=torch.bfloat16, device=x.device)
        _check(L.lib().ni_unet_avgpool2(0, N, _shape(D, H, W), C, _ptr(x), _ptr(y), L.stream_of(x)), "unet_avgpool2")
        return y

    def _pool_bwd(self, dy, shape, C):
        N, (D, H, W) = dy.shape[0], shape
        dx = torch.empty(N, D, H, W, C, dtype=torch.bfloat16, device=dy.device)
        _check(L.lib().ni_unet_avgpool2(1, N, _shape(D, H, W), C, _ptr(dy), _ptr(dx), L.stream_of(dy)),
               "unet_avgpool2")
        return dx

    def _upcat(self, a, skip, shape, C1, C2):
        N, (D, H, W) = a.shape[0], shape
        y = torch.empty(N, D, H, W, C1 + C2, dtype=torch.bfloat16, device=a.device)
        _check(L.lib().ni_unet_up2_concat(0, N, _shape(D, H, W), C1, C2, _ptr(a), _ptr(skip), _ptr(y),
                                          L.stream_of(a)), "unet_up2_concat")
        return y

    def _upcat_bwd(self, dy, shape, C1, C2):
        N, (D, H, W) = dy.shape[0], shape
        da = torch.empty(N, D // 2, H // 2, W // 2, C1, dtype=torch.bfloat16, device=dy.device)
        ds = torch.empty(N, D, H, W, C2, dtype=torch.bfloat16, device=dy.device)
        _check(L.lib().ni_unet_up2_concat(1, N, _shape(D, H, W), C1, C2, _ptr(da), _ptr(ds), _ptr(dy),
                                          L.stream_of(dy)), "unet_up2_concat")
        return da, ds

    def _lrelu_bwd(self, dy, y):
        out = torch.empty_like(dy)
        _check(L.lib().ni_unet_lrelu_bwd(dy.numel(), _ptr(dy), _ptr(y), _ptr(out), L.stream_of(dy)), "unet_lrelu_bwd")
        return out

    def _bwdw(self, x, dy, shape, ci, co, grads, name):
        N, (D, H, W) = x.shape[0], shape
        dw = torch.zeros(co, _kp(ci), dtype=torch.float32, device=x.device)
        db = torch.zeros(co, dtype=torch.float32, device=x.device)
        _check(L.lib().ni_unet_conv3_bwd_weight(N, _shape(D, H, W), ci, co, _ptr(x), _ptr(dy), _ptr(dw), _ptr(db),
                                                L.stream_of(x)), "unet_conv3_bwd_weight")
        grads[name] = (dw, db)

    # -- network --------------------------------------------------------------
    def pack(self, g, h, sg=1.0, sh=1.0):
        """Channels-last bf16 network input from channel-first fp32 g, h."""
        N, D, H, W = g.shape[0], *g.shape[2:]
        x = torch.empty(N, D, H, W, 16, dtype=torch.bfloat16, device=g.device)
        CC = 0 if h is None else h.shape[1]
        _check(L.lib().ni_unet_pack_input(N, D * H * W, _ptr(g), _ptr(h), CC, _ptr(x), float(sg), float(sh),
                                          L.stream_of(g)), "unet_pack_input")
        return x

    def forward(self, x, keep=False):
        """x: packed input [N, D, H, W, 16] bf16 -> fp32 [N, D, H, W, 3].
        keep=True also returns the activations needed by backward()."""
        N, D, H, W, _ = x.shape
        s0, s1, s2 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4)
        c = self._conv
        a1 = c(x, s0, 16, 16, self.wk["e1a"], self.bias["e1a"], 1)
        a2 = c(a1, s0, 16, 16, self.wk["e1b"], self.bias["e1b"], 1)
        p1 = self._pool(a2, s0, 16)
        a3 = c(p1, s1, 16, 32, self.wk["e2a"], self.bias["e2a"], 1)
        a4 = c(a3, s1, 32, 32, self.wk["e2b"], self.bias["e2b"], 1)
        p2 = self._pool(a4, s1, 32)
        a5 = c(p2, s2, 32, 64, self.wk["b1"], self.bias["b1"], 1)
        a6 = c(a5, s2, 64, 64, self.wk["b2"], self.bias["b2"], 1)
        c2 = self._upcat(a6, a4, s1, 64, 32)
        a7 = c(c2, s1, 96, 32, self.wk["d2a"], self.bias["d2a"], 1)
        a8 = c(a7, s1, 32, 32, self.wk["d2b"], self.bias["d2b"], 1)
        if keep:    # backward needs the concatenation itself
            c1 = self._upcat(a8, a2, s0, 32, 16)
            a9 = c(c1, s0, 48, 16, self.wk["d1a"], self.bias["d1a"], 1)
        else:
            c1 = None
            a9 = self._conv_upcat(a8, a2, s0, 32, 16, 16, self.wk["d1a"], self.bias["d1a"], 1)
        a10 = c(a9, s0, 16, 16, self.wk["d1b"], self.bias["d1b"], 1)
        out = c(a10, s0, 16, 16, self.wk["head"], self.bias["head"], 0, out_f32=True, cos=3)
        if keep:
            return out, dict(x=x, a1=a1, a2=a2, p1=p1, a3=a3, a4=a4, p2=p2, a5=a5, a6=a6, c2=c2, a7=a7, a8=a8,
                             c1=c1, a9=a9, a10=a10)
        return out

    def backward(self, dout, act):
        """dout: fp32 [N, D, H, W, 3] gradient of the output.  Returns
        {name: (dW [CO_s, KP] fp32, db [CO_s] fp32)} for every layer."""
        x = act["x"]
        N, D, H, W, _ = x.shape
        s0, s1, s2 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4)
        g = {}
        c = self._conv
        d = torch.zeros(N, D, H, W, 16, dtype=torch.bfloat16, device=x.device)
        d[..., :3] = dout.to(torch.bfloat16)
        self._bwdw(act["a10"], d, s0, 16, 16, g, "head")
        da = c(d, s0, 16, 16, self.wt["head"], None, 0)
        d = self._lrelu_bwd(da, act["a10"])
        self._bwdw(act["a9"], d, s0, 16, 16, g, "d1b")
        da = c(d, s0, 16, 16, self.wt["d1b"], None, 0)
        d = self._lrelu_bwd(da, act["a9"])
        self._bwdw(act["c1"], d, s0, 48, 16, g, "d1a")
        dc1 = c(d, s0, 16, 48, self.wt["d1a"], None, 0)
        da8, da2_skip = self._upcat_bwd(dc1, s0, 32, 16)
        d = self._lrelu_bwd(da8, act["a8"])
        self._bwdw(act["a7"], d, s1, 32, 32, g, "d2b")
        da = c(d, s1, 32, 32, self.wt["d2b"], None, 0)
        d = self._lrelu_bwd(da, act["a7"])
        self._bwdw(act["c2"], d, s1, 96, 32, g, "d2a")
        dc2 = c(d, s1, 32, 96, self.wt["d2a"], None, 0)
        da6, da4_skip = self._upcat_bwd(dc2, s1, 64, 32)
        d = self._lrelu_bwd(da6, act["a6"])
        self._bwdw(act["a5"], d, s2, 64, 64, g, "b2")
        da = c(d, s2, 64, 64, self.wt["b2"], None, 0)
        d = self._lrelu_bwd(da, act["a5"])
        self._bwdw(act["p2"], d, s2, 32, 64, g, "b1")
        dp2 = c(d, s2, 64, 32, self.wt["b1"], None, 0)
        da4 = da4_skip.float().add_(self._pool_bwd(dp2, s1, 32).float()).to(torch.bfloat16)
        d = self._lrelu_bwd(da4, act["a4"])
        self._bwdw(act["a3"], d, s1, 32, 32, g, "e2b")
        da = c(d, s1, 32, 32, self.wt["e2b"], None, 0)
        d = self._lrelu_bwd(da, act["a3"])
        self._bwdw(act["p1"], d, s1, 16, 32, g, "e2a")
        dp1 = c(d, s1, 32, 16, self.wt["e2a"], None, 0)
        da2 = da2_skip.float().add_(self._pool_bwd(dp1, s0, 16).float()).to(torch.bfloat16)
        d = self._lrelu_bwd(da2, act["a2"])
        self._bwdw(act["a1"], d, s0, 16, 16, g, "e1b")
        da = c(d, s0, 16, 16, self.wt["e1b"], None, 0)
        d = self._lrelu_bwd(da, act["a1"])
        self._bwdw(x, d, s0, 16, 16, g, "e1a")
        return g

    def grads_to_params(self, g):
        """Kernel-layout gradients -> gradients of the fp32 masters."""
        out = {}
        for name, ci, ci_s, co, co_s, act in LAYERS:
            dw, db = g[name]
            w = dw[:, :27 * ci_s].reshape(co_s, 27, ci_s)[:co, :, :ci].reshape(co, 3, 3, 3, ci)
            out[name + ".w"] = w
            out[name + ".b"] = db[:co]
        return out

    # -- solver-facing ----------------------------------------------------------
    def apply(self, r, h, penalty, voxel_size=(1, 1, 1), bound="dft", hscale=1.0):
        """x ~ (H + L)^-1 r for channel-first fp32 r [N, 3, D, H, W]: the
        block-Jacobi estimate (H + diag L)^-1 r plus the network's correction,
        evaluated on (r / rms(r), H * hscale) and rescaled by rms(r)."""
        N = r.shape[0]
        V = r[0, 0].numel()
        s = float(r.pow(2).mean().sqrt())
        s = s if s > 0 else 1.0
        out = block_jacobi(r, h, penalty, voxel_size, bound)
        y = self.forward(self.pack(r, h, 1.0 / s, hscale))
        _check(L.lib().ni_unet_unpack_output(N, V, _ptr(y), _ptr(out), float(s), 1, L.stream_of(r)),
               "unet_unpack_output")
        return out


class Adam:
    """Adam on the fp32 masters of a UNet (torch elementwise ops)."""

    def __init__(self, net, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.net, self.lr, self.b1, self.b2, self.eps, self.t = net, lr, betas[0], betas[1], eps, 0
        self.m = {k: torch.zeros_like(v) for k, v in net.params.items()}
        self.v = {k: torch.zeros_like(v) for k, v in net.params.items()}

    def step(self, grads, lr=None):
        lr = self.lr if lr is None else lr
        self.t += 1
        c1, c2 = 1 - self.b1 ** self.t, 1 - self.b2 ** self.t
        for k, g in grads.items():
            self.m[k].mul_(self.b1).add_(g, alpha=1 - self.b1)
            self.v[k].mul_(self.b2).addcmul_(g, g, value=1 - self.b2)
            self.net.params[k].addcdiv_(self.m[k] / c1, (self.v[k] / c2).sqrt_().add_(self.eps), value=-lr)
        self.net.refresh()


def block_jacobi(r, h, penalty, voxel_size=(1, 1, 1), bound="dft"):
    """(H + diag L)^-1 r voxel by voxel (the reference's precond_grid)."""
    return SV.precond_grid_cf(h, r, penalty, voxel_size, bound)


def train_step(net, opt, g, h, penalty, x_target, lr=None, voxel_size=(1, 1, 1), bound="dft"):
    """One Adam step on one system: the network corrects the block-Jacobi
    estimate of x = (H + L)^-1 g; loss = ||estimate - x||^2 / ||x||^2.
    Returns (loss before the step, gradients)."""
    s = float(g.pow(2).mean().sqrt())
    base = block_jacobi(g, h, penalty, voxel_size, bound)
    target = (x_target - base).permute(0, 2, 3, 4, 1) / s
    out, act = net.forward(net.pack(g, h, 1.0 / s), keep=True)
    err = out - target
    den = float((x_target / s).pow(2).sum())
    loss = float(err.pow(2).sum()) / den
    grads = net.grads_to_params(net.backward(err * (2.0 / den), act))
    opt.step(grads, lr=lr)
    return loss, grads


# ---------------------------------------------------------------------------
# learned-preconditioned conjugate gradient
# ---------------------------------------------------------------------------
def _dot(a, b):
    return torch.dot(a.reshape(-1), b.reshape(-1))


def pcg_learned(net, hessian, gradient, penalty, voxel_size=(1, 1, 1), bound="dft", tol=1e-4,
                max_iter=50, stall=0.9, fallback_cycles=20, hscale=1.0, warm_start=True):
    """Flexible PCG on (H + L) x = g with the U-Net as warm start and
    preconditioner (Polak-Ribiere beta, as the preconditioner is nonlinear).
    Channel-first contiguous fp32 tensors.  When the residual ratio of an
    iteration exceeds `stall` twice in a row, FMG V-cycles take over from
    the current iterate.  A non-finite value anywhere in the recurrence (a
    network output with NaN / Inf, an overflowing step) counts as a stall:
    the iterate is never updated with a non-finite step, and when the
    network's warm start itself is not finite the fallback starts from zero.
    Returns (x, relative residual, iterations, used_fallback)."""
    g, h = gradient, hessian
    gn = float(g.pow(2).sum().sqrt())
    it = 0
    x = net.apply(g, h, penalty, voxel_size, bound, hscale) if warm_start else torch.zeros_like(g)
    r = SV.residual_grid_cf(x, g, h, penalty, voxel_size, bound)
    rn = float(r.pow(2).sum().sqrt())
    fresh = not math.isfinite(rn) or not warm_start   # no usable iterate yet
    if not math.isfinite(rn):
        x = torch.zeros_like(g)
    elif rn <= tol * gn:
        return x, rn / gn, 0, False
    else:
        z = net.apply(r, h, penalty, voxel_size, bound, hscale)
        p = z.clone()
        rz = _dot(r, z)
        Ap = torch.empty_like(g)
        slow = 0
        for it in range(1, max_iter + 1):
            SV.regulariser_grid_cf(p, h, penalty, voxel_size, bound, out=Ap)
            pap, rzf = float(_dot(p, Ap)), float(rz)
            if not (math.isfinite(pap) and math.isfinite(rzf)):
                break                      # the direction is unusable; x is untouched
            alpha = rzf / max(pap, 1e-30)
            x.add_(p, alpha=alpha)
            r.sub_(Ap, alpha=alpha)
            fresh = False
            rn_new = float(r.pow(2).sum().sqrt())
            if not math.isfinite(rn_new):
                x.sub_(p, alpha=alpha)     # back to the last finite iterate
                fresh = not warm_start and it == 1
                break
            if rn_new <= tol * gn:
                return x, rn_new / gn, it, False
            slow = slow + 1 if rn_new > stall * rn else 0
            rn = rn_new
            if slow >= 2:
                break
            z_new = net.apply(r, h, penalty, voxel_size, bound, hscale)
            rz_new = _dot(r, z_new)
            beta = (rz_new - _dot(r, z)) / rz.clamp_min(1e-30)
            rz = rz_new
            z = z_new
            p.mul_(beta.clamp_min(0)).add_(z)
        if not bool(torch.isfinite(x).all()):
            fresh = True
            x = torch.zeros_like(g)
    # fallback: FMG V-cycles from the current iterate (their own residual test
    # decides convergence); a full FMG pass first when starting from zero
    x2, rel, ncyc = SV.solve_to_tolerance(h, g, penalty, voxel_size, bound, tol=tol, nb_cycles=2 if fresh else 0,
                                          nb_iter=4,
                                          max_vcycles=fallback_cycles, max_levels=5, x=x, use_cg=True)
    return x2, rel, it, True


# ---------------------------------------------------------------------------
# training data and training loop
# ---------------------------------------------------------------------------
BASE_PENALTY = (1e-4, 1e-3, 0.2, 0.05, 0.2)   # absolute, membrane, bending, lame shear, lame div


def _smooth_gpu(field, sigma):
    """Periodic Gaussian smoothing over the last 3 dims (FFT, on the device)."""
    shape = field.shape[-3:]
    freqs = [torch.fft.fftfreq(n, device=field.device, dtype=torch.float32) for n in shape[:-1]]
    freqs.append(torch.fft.rfftfreq(shape[-1], device=field.device, dtype=torch.float32))
    k2 = sum(f.reshape([-1 if i == j else 1 for j in range(3)]) ** 2 for i, f in enumerate(freqs))
    H = torch.exp(-2 * (math.pi ** 2) * sigma ** 2 * k2)
    return torch.fft.irfftn(torch.fft.rfftn(field, s=shape, dim=(-3, -2, -1)) * H, s=shape, dim=(-3, -2, -1))


def make_system(shape, seed, device="cuda", sigma_g=2.0, spread=(0.3, 3.0)):
    """One Gauss-Newton system of the fmg family: g (smoothed N(0,1), unit
    std), H (diagonal U[0.5, 2], off-diagonal U[-0.2, 0.2]) channel-first
    fp32, and the base penalties each scaled by a log-uniform factor in
    `spread`.  Returns (g [1,3,*shape], h [1,6,*shape], penalty tuple)."""
    gen = torch.Generator(device=device).manual_seed(int(seed))
    g = _smooth_gpu(torch.randn((1, 3) + tuple(shape), generator=gen, device=device), sigma_g)
    g = g / g.std()
    hd = torch.rand((1, 3) + tuple(shape), generator=gen, device=device) * 1.5 + 0.5
    ho = torch.rand((1, 3) + tuple(shape), generator=gen, device=device) * 0.4 - 0.2
    h = torch.cat([hd, ho], 1).contiguous()
    lo, hi = math.log(spread[0]), math.log(spread[1])
    f = torch.exp(torch.rand(5, generator=gen, device=device) * (hi - lo) + lo).tolist()
    pen = tuple(b * s for b, s in zip(BASE_PENALTY, f))
    return g.contiguous(), h, pen


def train(net=None, shape=(96, 96, 96), n_systems=2000, epochs=1, lr=1e-3, seed=0, log=None,
          solve_tol=1e-5, device="cuda"):
    """Train the U-Net on (g, H) -> FMG-solution pairs (relative squared error
    of the rms(g)-normalised solution).  Each system is generated and solved
    on the fly (FMG with the PCG smoother to `solve_tol`), then used for one
    Adam step per epoch.  Returns the network and the loss history."""
    net = UNet(device, seed=seed) if net is None else net
    opt = Adam(net, lr=lr)
    hist = []
    total = n_systems * epochs
    step = 0
    cache = {}
    for ep in range(epochs):
        for i in range(n_systems):
            if i in cache:
                g, h, pen, x = cache[i]
            else:
                g, h, pen = make_system(shape, 1000003 * (seed + 1) + i, device)
                x, rel, _ = SV.solve_to_tolerance(h, g, pen, tol=solve_tol, nb_cycles=2, nb_iter=4, max_levels=5,
                                                  max_vcycles=40, use_cg=True)
                if epochs > 1:
                    cache[i] = (g, h, pen, x)
            cur = lr * 0.5 * (1 + math.cos(math.pi * step / total))   # cosine decay
            loss, _ = train_step(net, opt, g, h, pen, x, lr=cur)
            hist.append(loss)
            step += 1
            if log is not None and (step % 50 == 0 or step == 1):
                log(f"step {step}/{total} loss {loss:.4f} (mean of last 50: {np.mean(hist[-50:]):.4f})")
    return net, hist


def default_weights_path():
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "precond_unet.npz")


if __name__ == "__main__":
    import argparse
    import os
    import time
    ap = argparse.ArgumentParser(description="train the learned preconditioner's U-Net on the GPU")
    ap.add_argument("--out", default=default_weights_path())
    ap.add_argument("--systems", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--size", type=int, default=96)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    t0 = time.time()
    net, hist = train(shape=(a.size,) * 3, n_systems=a.systems, epochs=a.epochs, lr=a.lr, seed=a.seed,
                      log=lambda m: print(f"[{time.time() - t0:7.1f}s] {m}", flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    print(f"saved {a.out}; final mean loss {np.mean(hist[-100:]):.4f}; {time.time() - t0:.1f}s", flush=True)
