"""GPU parity of the solver kernels with voxel-wise RLS weights (absolute /
membrane energies) against the oracle, itself pinned by the reference's
compiled solver (tests/test_solver_rls_oracle.py).

Tolerances (norm-relative, max|a-b| / max|b|): FP64 1e-10 for the
operators, precond and PCG, 1e-9 for relax / FMG (several sweeps); FP32
2e-5 for single operators, 1e-4 for relax / FMG.  Batches of 2 with
different weights per batch element.
"""
import numpy as np
import pytest
import torch

from oracle import solver as S

pytestmark = pytest.mark.gpu
DEV = "cuda"
PENS = [(1e-3, 0.0), (1e-4, 0.1), (0.0, 0.05), (0.02, 0.3)]


def rel(a, b):
    a = a.detach().cpu().double().numpy() if torch.is_tensor(a) else a
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def T(a, dt):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dt)


def system(rng, shape, CC, N=2):
    g = rng.standard_normal((N, 3) + shape)
    x = rng.standard_normal((N, 3) + shape)
    w = rng.uniform(0.2, 2.0, (N, 1) + shape)
    h = None
    if CC:
        h = rng.uniform(0.5, 2, (N, CC) + shape)
        if CC == 6:
            h[:, 3:] = rng.uniform(-0.2, 0.2, (N, 3) + shape)
    return g, x, w, h


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
@pytest.mark.parametrize("bound", ["dft", "dct2", "replicate", "dst2", "zero"])
def test_rls_operators_vs_oracle(dt, bound):
    from nitorch_amd import solve as SV
    rng = np.random.default_rng(71)
    tol_op = 1e-10 if dt == torch.float64 else 2e-5
    tol_it = 1e-9 if dt == torch.float64 else 1e-4
    for k, (a, m) in enumerate(PENS):
        CC = (0, 1, 3, 6)[k % 4]
        shape = (9, 7, 10) if k % 2 else (6, 11, 5)
        g, x, w, h = system(rng, shape, CC)
        vs = (1.0, 1.3, 0.8)
        pen = (a, m, 0, 0, 0)
        kw = dict(absolute=a, membrane=m, voxel_size=vs, bound=bound, weight=w)
        case = (bound, a, m, CC)
        out = SV.regulariser_grid_cf(T(x, dt), T(h, dt), pen, vs, bound, weight=T(w, dt))
        assert rel(out, S.regulariser_grid(x, h, **kw)) <= tol_op, case
        out = SV.precond_grid_cf(T(h, dt), T(g, dt), pen, vs, bound, weight=T(w, dt))
        assert rel(out, S.precond_grid(g, h, **kw)) <= tol_op, case
        xo, it = SV.pcg_grid_cf(T(h, dt), T(g, dt), T(x * 0.1, dt), pen, vs, bound, nb_iter=5, weight=T(w, dt))
        ref, _ = S.pcg_grid(h, g, x * 0.1, nb_iter=5, **kw)
        assert it == 5 and rel(xo, ref) <= tol_it, case
        if m:
            xr = SV.relax_grid_cf(T(g, dt), T(x, dt), T(h, dt), pen, vs, bound, nb_iter=3, weight=T(w, dt))
            ref = S.relax_grid(g, h, x, nb_iter=3, product_order=True, **kw)
            assert rel(xr, ref) <= tol_it, case


@pytest.mark.parametrize("use_cg", [False, True])
def test_rls_fmg_vs_oracle(use_cg):
    """Weighted FMG (weights restricted with the gradients, never rescaled)
    against the oracle, and the weighted residual it leaves, on a level-4
    pyramid; the large case also runs the single-work-group coarse relax."""
    from nitorch_amd import solve as SV
    rng = np.random.default_rng(72)
    for shape, dt, tol in (((12, 10, 9), torch.float64, 1e-9), ((40, 36, 34), torch.float32, 1e-4)):
        g, _, w, h = system(rng, shape, 6)
        pen = (1e-3, 0.3, 0, 0, 0)
        x = SV.fmg_grid_cf(T(h, dt), T(g, dt), None, pen, (1, 1, 1), "dft", nb_cycles=2, nb_iter=2,
                           use_cg=use_cg, weight=T(w, dt))
        ref = S.fmg_grid(h, g, absolute=1e-3, membrane=0.3, bound="dft", nb_cycles=2, nb_iter=2,
                         product_order=True, use_cg=use_cg, weight=w)
        assert rel(x, ref) <= tol, (shape, use_cg)
    # solve to tolerance with weights: residual of the weighted system
    g, _, w, h = system(rng, (48, 40, 36), 6, N=1)
    gt, ht, wt = T(g, torch.float32), T(h, torch.float32), T(w, torch.float32)
    x, relr, ncyc = SV.solve_to_tolerance(ht, gt, (1e-3, 0.3, 0, 0, 0), tol=1e-4, use_cg=use_cg, weight=wt)
    assert relr <= 1e-4
    res = g - S.regulariser_grid(x.double().cpu().numpy(), h, absolute=1e-3, membrane=0.3, weight=w)
    assert np.linalg.norm(res) / np.linalg.norm(g) <= 1.2e-4


def test_rls_reference_layer_and_errors():
    """The channel-last wrappers take (N, *spatial) weights; bending or Lame
    penalties with weights are an error, as in the reference."""
    from nitorch_amd import solve as SV
    rng = np.random.default_rng(73)
    g, x, w, h = system(rng, (8, 9, 10), 6, N=1)
    gl = T(np.moveaxis(g, 1, -1), torch.float64)
    hl = T(np.moveaxis(h, 1, -1), torch.float64)
    wl = T(w[:, 0], torch.float64)
    out = SV.regulariser_grid(T(np.moveaxis(x, 1, -1), torch.float64), wl, hl, absolute=1e-3, membrane=0.2)
    ref = S.regulariser_grid(x, h, absolute=1e-3, membrane=0.2, weight=w)
    assert rel(out, np.moveaxis(ref, 1, -1)) <= 1e-10
    sol = SV.fmg_grid(hl, gl, wl, absolute=1e-3, membrane=0.2, solver="relax", nb_cycles=2, nb_iter=2)
    ref = S.fmg_grid(h, g, absolute=1e-3, membrane=0.2, nb_cycles=2, nb_iter=2, product_order=True, weight=w)
    assert rel(sol, np.moveaxis(ref, 1, -1)) <= 1e-9
    with pytest.raises(RuntimeError):
        SV.regulariser_grid(gl, wl, hl, absolute=1e-3, bending=0.1)
