"""What the trained preconditioner gains over the method without it, on a
held-out system of the training family (64^3, seed outside the training
range): its one-shot estimate of x = (H + L)^-1 g is much closer than the
block-Jacobi estimate the network corrects, and the flexible PCG needs fewer
iterations to 1e-4 than the same PCG with the untrained network (which is
block-Jacobi, its last layer being zero).  The solution is checked with the
FP64 oracle operator."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module")
def setup():
    from nitorch_amd import learned as LN
    from nitorch_amd import solve as SV
    path = LN.default_weights_path()
    if not os.path.exists(path):
        pytest.skip("no trained weights")
    g, h, pen = LN.make_system((64, 64, 64), seed=987654321, device=DEV)
    xs, rel, _ = SV.solve_to_tolerance(h, g, pen, tol=1e-6, nb_cycles=2, nb_iter=4, max_levels=5,
                                       max_vcycles=60, use_cg=True)
    assert rel <= 2e-6
    return LN, LN.UNet.load(path, DEV), g, h, pen, xs


def test_one_shot_estimate(setup):
    LN, net, g, h, pen, xs = setup
    est = net.apply(g, h, pen)
    bj = LN.block_jacobi(g, h, pen)
    e_net = float((est - xs).norm() / xs.norm())
    e_bj = float((bj - xs).norm() / xs.norm())
    assert e_net < 0.5 * e_bj, (e_net, e_bj)


def test_pcg_iterations_gain(setup):
    from oracle import solver as S
    LN, net, g, h, pen, xs = setup
    x1, r1, it1, fb1 = LN.pcg_learned(net, h, g, pen, tol=1e-4, max_iter=300)
    x0, r0, it0, fb0 = LN.pcg_learned(LN.UNet(DEV), h, g, pen, tol=1e-4, max_iter=300)
    assert r1 <= 1e-4 and r0 <= 1e-4
    assert not fb1
    assert it1 < it0, (it1, it0)
    # FP64 residual of the learned solution
    res = g.double().cpu().numpy() - S.regulariser_grid(x1.double().cpu().numpy(), h.double().cpu().numpy(),
                                                         absolute=pen[0], membrane=pen[1], bending=pen[2],
                                                         lame=(pen[3], pen[4]))
    assert np.linalg.norm(res) / float(g.norm()) <= 2e-4


class _StubNet:
    """A preconditioner whose output is unusable: NaN, zero or random."""

    def __init__(self, kind, seed=0):
        self.kind = kind
        self.gen = torch.Generator(device=DEV).manual_seed(seed)

    def apply(self, r, h, penalty, voxel_size=(1, 1, 1), bound="dft", hscale=1.0):
        if self.kind == "nan":
            return torch.full_like(r, float("nan"))
        if self.kind == "inf":
            out = torch.zeros_like(r)
            out[0, 0, 0, 0, 0] = float("inf")
            return out
        if self.kind == "zeros":
            return torch.zeros_like(r)
        return torch.randn(r.shape, generator=self.gen, device=r.device) * 10.0


def _fp64_rel_residual(x, g, h, pen):
    from oracle import solver as S
    res = g.double().cpu().numpy() - S.regulariser_grid(x.double().cpu().numpy(), h.double().cpu().numpy(),
                                                         absolute=pen[0], membrane=pen[1], bending=pen[2],
                                                         lame=(pen[3], pen[4]))
    return np.linalg.norm(res) / float(g.double().norm())


@pytest.mark.parametrize("kind", ["nan", "inf", "zeros", "garbage"])
@pytest.mark.parametrize("warm_start", [True, False])
def test_fallback_on_bad_network(kind, warm_start):
    """A network that returns NaN, Inf, zeros or garbage is caught by the
    stall test, and the FMG fallback still reaches the tolerance (checked with
    the FP64 oracle operator)."""
    from nitorch_amd import learned as LN
    g, h, pen = LN.make_system((32, 32, 32), seed=424242, device=DEV)
    print(f"bad network {kind} warm_start={warm_start}", flush=True)
    x, rel, it, fb = LN.pcg_learned(_StubNet(kind), h, g, pen, tol=1e-4, max_iter=20, warm_start=warm_start)
    assert fb
    assert bool(torch.isfinite(x).all())
    assert rel <= 1e-4
    assert _fp64_rel_residual(x, g, h, pen) <= 2e-4


def test_forced_fallback_real_network(setup):
    """The trained network with the stall threshold at 0 (every iteration
    counts as stalled): the FMG fallback finishes from the PCG iterate."""
    LN, net, g, h, pen, xs = setup
    x, rel, it, fb = LN.pcg_learned(net, h, g, pen, tol=1e-4, max_iter=50, stall=0.0)
    assert fb and it == 2
    assert rel <= 1e-4
    assert _fp64_rel_residual(x, g, h, pen) <= 2e-4
    x, rel, it, fb = LN.pcg_learned(net, h, g, pen, tol=1e-4, max_iter=1)
    assert fb and it == 1 and rel <= 1e-4
    assert _fp64_rel_residual(x, g, h, pen) <= 2e-4
