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
