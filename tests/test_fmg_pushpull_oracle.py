"""The oracle of the pull/push multigrid (oracle/fmg_pushpull.py) against
closed forms: the transfer grid, partition of unity of the quadratic
prolongation, constants preserved by the normalised restriction, the
adjoint pair (restriction = scaled push of order 1), the one-level case
(= the level smoother alone) and convergence to the CG solution."""
import numpy as np
import pytest

from oracle import fmg_pushpull as F
from oracle import pushpull as PP
from oracle import solver as S

PEN = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2))


def _system(shape, seed, cc=6):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((1, 3) + shape)
    h = rng.uniform(0.5, 2, (1, cc) + shape)
    if cc == 6:
        h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
    return g, h


def test_make_grid_closed_form():
    g = F.make_grid((8, 6, 4), (4, 3, 2))
    i = np.arange(8)
    assert np.allclose(g[:, 0, 0, 0], (i + 0.5) / 2 - 0.5)
    assert np.allclose(g[0, :, 0, 1], (np.arange(6) + 0.5) / 2 - 0.5)
    g = F.make_grid((7, 5, 3), (4, 3, 2))   # odd sizes: ratio n_c / n_f
    assert np.allclose(g[:, 0, 0, 0], np.arange(7) * 4 / 7 + 0.5 * (4 / 7 - 1))


@pytest.mark.parametrize("bound", ["dft", "dct2"])
def test_prolong_restrict_constants(bound):
    c = np.full((2, 3, 5, 4, 3), 1.7)
    assert np.allclose(F.prolong(c, (10, 8, 6), bound), 1.7, atol=1e-12)
    f = np.full((2, 3, 10, 8, 6), -0.4)
    assert np.allclose(F.restrict(f, (5, 4, 3), bound), -0.4, atol=1e-12)


def test_restrict_is_scaled_adjoint_of_linear_pull():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 2, 10, 8, 6))
    y = rng.standard_normal((1, 2, 5, 4, 3))
    grid = F.make_grid((10, 8, 6), (5, 4, 3))[None]
    lhs = np.sum(F.restrict(x, (5, 4, 3), "dft") * y)
    rhs = (5 * 4 * 3) / (10 * 8 * 6) * np.sum(x * PP.grid_pull(y, grid, "dft", 1, 1))
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_single_level_is_the_smoother():
    g, h = _system((6, 5, 4), 1)
    x = F.fmg(h, g, max_levels=0, nb_iter=3, **PEN)
    ref = S.pcg_grid(h, g, None, nb_iter=3, tol=0.0, **PEN)[0]
    assert np.abs(x - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("optim", ["cg", "relax"])
def test_converges_to_cg_solution(optim):
    g, h = _system((12, 10, 8), 2)
    xs = S.cg_grid(h, g, tol=1e-12, **PEN)[0]
    x = F.fmg(h, g, nb_cycles=3, nb_iter=4, optim=optim, **PEN)
    bj = S.precond_grid(g, h, **PEN)
    err = np.linalg.norm(x - xs) / np.linalg.norm(xs)
    assert err < 0.1 * np.linalg.norm(bj - xs) / np.linalg.norm(xs), err
    res = np.linalg.norm(g - S.regulariser_grid(x, h, **PEN)) / np.linalg.norm(g)
    assert res < 5e-2, res


@pytest.mark.parametrize("shape", [(16, 12, 10), (13, 9, 7)])
@pytest.mark.parametrize("bound", ["dft", "dct2"])
@pytest.mark.parametrize("optim", ["cg", "relax"])
def test_equals_compiled_multigrid(shape, bound, optim):
    """The pull/push multigrid and the compiled one (quadratic prolongation,
    linear restriction, channel rescaling) are the same algorithm: their
    iterates agree to rounding, odd sizes included."""
    g, h = _system(shape, 3)
    x1 = F.fmg(h, g, nb_cycles=2, nb_iter=2, optim=optim, bound=bound, **PEN)
    x2 = S.fmg_grid(h, g, nb_cycles=2, nb_iter=2, use_cg=optim == "cg", product_order=True,
                    bound=bound, **PEN)
    assert np.abs(x1 - x2).max() <= 1e-12 * np.abs(x2).max()
