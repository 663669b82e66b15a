"""CPU tests of the solver oracle with voxel-wise RLS weights (absolute and
membrane energies, nitorch/_C/src/impl/regulariser_grid_common.cpp:1171-1248,
relax_grid_common.cpp:1415-1500, precond_grid_common.cpp:800-840,
fmg_common.cpp:176-181).

Pinned by the reference's own vectors (tests/golden/solver_rls_golden.npz,
written by tests/golden/make_rls_golden.py from oracle/_ref) and by
properties that a dropped or misplaced term breaks: a constant weight map
equals the unweighted operator scaled by that constant; the weighted
operator is symmetric (face weights are shared by both neighbours); the
relax diagonal is the diagonal of the operator; relax and FMG converge to
the CG solution of the weighted system.
"""
import os

import numpy as np
import pytest

from oracle import solver as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "solver_rls_golden.npz")
BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module")
def golden():
    g = np.load(GOLDEN)
    keys = sorted({k.split("_")[0] for k in g.files})
    return {k: {f.split("_", 1)[1]: g[f] for f in g.files if f.startswith(k + "_")} for k in keys}


def test_rls_operators_golden(golden):
    n = 0
    for key, c in golden.items():
        if not key.startswith("w"):
            continue
        b, CC = [int(v) for v in c["meta"]]
        a, m, be, mu, la, *vs = [float(v) for v in c["pen"]]
        kw = dict(absolute=a, membrane=m, voxel_size=vs, bound=BOUNDS[b], weight=c["w"])
        h = c.get("h")
        assert rel(S.regulariser_grid(c["x"], h, **kw), c["reg"]) < 1e-12, key
        if "relax" in c:
            assert rel(S.relax_grid(c["g"], h, c["x"], nb_iter=2, **kw), c["relax"]) < 1e-12, key
        assert rel(S.precond_grid(c["g"], h, **kw), c["precond"]) < 1e-12, key
        x, it = S.pcg_grid(h, c["g"], c["x"] * 0.1, nb_iter=5, tol=0.0, **kw)
        assert it == 5 and rel(x, c["pcg"]) < 1e-10, key
        n += 1
    assert n == 20


def test_rls_fmg_golden(golden):
    n = 0
    for key, c in golden.items():
        if not key.startswith("f"):
            continue
        b, use_cg = [int(v) for v in c["meta"]]
        a, m, be, mu, la = [float(v) for v in c["pen"]]
        out = S.fmg_grid(c["h"], c["g"], absolute=a, membrane=m, bound=BOUNDS[b], nb_cycles=2, nb_iter=2,
                         quirk_zscale=True, use_cg=bool(use_cg), weight=c["w"])
        assert rel(out, c["sol"]) < 1e-10, key
        n += 1
    assert n == 12


@pytest.mark.parametrize("bound", ["dft", "dct2", "replicate", "dct1"])
def test_constant_weight_scales_the_penalty(bound):
    """w = k everywhere: the weighted operator is the unweighted one with
    absolute and membrane multiplied by k (H untouched).  Bounds whose
    outside values carry a sign (zero, dst) also sign the neighbour's weight,
    as the reference reads weights with the field's rule, so they are not
    covered by this identity (the golden vectors cover them)."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 5, 6, 4))
    h = rng.uniform(0.5, 2, (2, 6, 5, 6, 4))
    h[:, 3:] *= 0.1
    k = 1.7
    w = np.full((2, 1, 5, 6, 4), k)
    out = S.regulariser_grid(x, h, absolute=0.3, membrane=0.2, voxel_size=[1.0, 1.2, 0.9], bound=bound, weight=w)
    ref = S.regulariser_grid(x, h, absolute=0.3 * k, membrane=0.2 * k, voxel_size=[1.0, 1.2, 0.9], bound=bound)
    assert rel(out, ref) < 1e-12


@pytest.mark.parametrize("bound", ["dft", "dct2", "replicate"])
def test_weighted_operator_is_symmetric_with_matching_diagonal(bound):
    """<u, L_w v> = <L_w u, v>, and the diagonal used by relax / precond is
    the diagonal of L_w (checked on unit vectors; away from the faces for the
    reflecting bounds, where a neighbour that folds onto the voxel itself
    drops out of the operator but not out of the reference's diagonal)."""
    rng = np.random.default_rng(2)
    shape = (1, 3, 4, 5, 3)
    w = rng.uniform(0.2, 2.0, (1, 1) + shape[2:])
    kw = dict(absolute=0.05, membrane=0.4, voxel_size=[1.0, 1.3, 0.8], bound=bound, weight=w)
    u, v = rng.standard_normal(shape), rng.standard_normal(shape)
    lu, lv = S.regulariser_grid(u, **kw), S.regulariser_grid(v, **kw)
    assert abs((u * lv).sum() - (lu * v).sum()) < 1e-12 * abs((u * lv).sum())
    # precond with an all-zero Hessian returns g / diag: compare with <e_i, L e_i>
    g = np.ones(shape)
    inv_diag = S.precond_grid(g, None, **kw)
    points = [(0, 1, 2, 3, 1), (0, 2, 1, 2, 1), (0, 0, 2, 1, 1)]
    if bound == "dft":
        points += [(0, 0, 0, 0, 0), (0, 2, 3, 4, 2)]
    for idx in points:
        e = np.zeros(shape)
        e[idx] = 1.0
        assert abs(S.regulariser_grid(e, **kw)[idx] * inv_diag[idx] - 1.0) < 1e-12, idx


@pytest.mark.parametrize("use_cg", [False, True])
def test_rls_relax_and_fmg_converge_to_cg(use_cg):
    rng = np.random.default_rng(3)
    shape = (1, 3, 12, 10, 9)
    g = rng.standard_normal(shape)
    h = rng.uniform(0.5, 2, (1, 6) + shape[2:])
    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape[2:])
    w = rng.uniform(0.2, 2.0, (1, 1) + shape[2:])
    kw = dict(absolute=1e-3, membrane=0.3, bound="dft", weight=w)
    xs, _ = S.cg_grid(h, g, tol=1e-13, **kw)
    x = None
    for _ in range(8):
        x = S.fmg_grid(h, g, x, nb_cycles=2, nb_iter=4, use_cg=use_cg, product_order=True, **kw)
    assert rel(x, xs) < 1e-6
    xr = S.relax_grid(g, h, None, nb_iter=400, product_order=True, **kw)
    assert rel(xr, xs) < 1e-6
