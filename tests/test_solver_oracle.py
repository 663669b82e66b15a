"""CPU tests of the solver oracle (FP64 restatement of regulariser_grid,
relax_grid, FMG transfers, precond_grid, pcg_grid and fmg_grid).

Pinned by the reference's own vectors (tests/golden/solver_golden.npz, written
by tests/golden/make_solver_golden.py from oracle/_ref) and by closed forms:
plane-wave eigenvalues of the scalar penalties under circulant (dft) bounds,
symmetry and positivity of (H + L), constant preservation of the transfer
operators, and convergence of relax / FMG to the CG solution.
"""
import os

import numpy as np
import pytest

from oracle import solver as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "solver_golden.npz")
BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.fixture(scope="module")
def golden():
    g = np.load(GOLDEN)
    keys = sorted({k.split("_")[0] for k in g.files})
    return {k: {f.split("_", 1)[1]: g[f] for f in g.files if f.startswith(k + "_")} for k in keys}


def test_regulariser_and_relax_golden(golden):
    n = 0
    for key, c in golden.items():
        if not key.startswith("r"):
            continue
        b, CC = [int(v) for v in c["meta"]]
        a, m, be, mu, la, *vs = [float(v) for v in c["pen"]]
        kw = dict(absolute=a, membrane=m, bending=be, lame=(mu, la), voxel_size=vs, bound=BOUNDS[b])
        h = c.get("h")
        out = S.regulariser_grid(c["x"], h, **kw)
        assert rel(out, c["reg"]) < 1e-12, key
        if "relax" in c:
            out = S.relax_grid(c["g"], h, c["x"], nb_iter=2, quirk_fold=True, **kw)
            assert rel(out, c["relax"]) < 1e-12, key
        n += 1
    assert n >= 40


def test_transfers_golden(golden):
    for key, c in golden.items():
        if not key.startswith("t"):
            continue
        b = BOUNDS[int(c["meta"][0])]
        assert rel(S.restrict(c["fine"], c["restrict"].shape[2:], b), c["restrict"]) < 1e-13, key
        assert rel(S.prolong(c["coarse"], c["prolong"].shape[2:], b), c["prolong"]) < 1e-13, key


def test_fmg_golden(golden):
    n = 0
    for key, c in golden.items():
        if not key.startswith("f"):
            continue
        b = BOUNDS[int(c["meta"][0])]
        a, m, be, mu, la = [float(v) for v in c["pen"]]
        kw = dict(absolute=a, membrane=m, bending=be, lame=(mu, la), bound=b, nb_cycles=2, nb_iter=2)
        if "sol" in c:
            out = S.fmg_grid(c["h"], c["g"], quirk_fold=True, quirk_zscale=True, **kw)
            assert rel(out, c["sol"]) < 1e-10, key
        else:
            out = S.fmg_grid(c["h"], c["g"], quirk_zscale=True, use_cg=True, **kw)
            assert rel(out, c["solcg"]) < 1e-10, key
        n += 1
    assert n == 12


def test_precond_and_pcg_golden(golden):
    n = 0
    for key, c in golden.items():
        if not key.startswith("c"):
            continue
        b, CC = [int(v) for v in c["meta"]]
        a, m, be, mu, la, *vs = [float(v) for v in c["pen"]]
        kw = dict(absolute=a, membrane=m, bending=be, lame=(mu, la), voxel_size=vs)
        h = c.get("h")
        assert rel(S.precond_grid(c["g"], h, **kw), c["precond"]) < 1e-13, key
        x, it = S.pcg_grid(h, c["g"], c["x"], bound=BOUNDS[b], nb_iter=5, tol=0.0, **kw)
        assert it == 5 and rel(x, c["pcg"]) < 1e-11, key
        if BOUNDS[b] == "dft" and h is None:
            # circulant L with no Hessian: (near-)singular system, the early-stop
            # iterate is rounding-dominated on both sides
            n += 1
            continue
        x, it = S.pcg_grid(h, c["g"], c["x"], bound=BOUNDS[b], nb_iter=50, tol=1e-3, **kw)
        assert rel(x, c["pcgtol"]) < 1e-11, key
        # the stopping iteration matters: one more or one fewer is far off
        x1, _ = S.pcg_grid(h, c["g"], c["x"], bound=BOUNDS[b], nb_iter=it + 1, tol=0.0, **kw)
        assert it == 1 or rel(x1, c["pcgtol"]) > 1e-6, key
        n += 1
    assert n == 30


def test_precond_is_block_diagonal_inverse():
    """(H + diag L) applied voxel-wise to precond(g) gives g back (up to the
    1e-6 diagonal inflation), for every Hessian layout."""
    rng = np.random.default_rng(11)
    shape = (4, 5, 3)
    kw = dict(absolute=0.2, membrane=0.1, bending=0.05, lame=(0.3, 0.4), voxel_size=(1.0, 1.2, 0.9))
    g = rng.standard_normal((2, 3) + shape)
    from oracle.solver import _setup  # noqa: F401  (library load)
    vx = 1 / np.array(kw["voxel_size"]) ** 2
    l0, l1, l2, mu, la = 0.2, 0.1, 0.05, 0.3, 0.4
    w000 = l2 * (6 * (vx ** 2).sum() + 8 * (vx[0] * vx[1] + vx[0] * vx[2] + vx[1] * vx[2])) + l1 * 2 * vx.sum() + l0
    diag = np.array([2 * mu * (vx.sum() + vx[d]) / vx[d] + 2 * la + w000 / vx[d] for d in range(3)])
    for CC in (0, 1, 3, 6):
        h = None
        Hm = np.zeros((2,) + shape + (3, 3))
        if CC:
            h = rng.uniform(0.5, 2, (2, CC) + shape)
            if CC == 6:
                h[:, 3:] = rng.uniform(-0.2, 0.2, (2, 3) + shape)
            hh = np.moveaxis(h, 1, -1)
            pairs = {1: [(0, 0, 0), (0, 1, 1), (0, 2, 2)], 3: [(0, 0, 0), (1, 1, 1), (2, 2, 2)],
                     6: [(0, 0, 0), (1, 1, 1), (2, 2, 2), (3, 0, 1), (4, 0, 2), (5, 1, 2)]}[CC]
            for k, i, j in pairs:
                Hm[..., i, j] = hh[..., k]
                Hm[..., j, i] = hh[..., k]
        A = Hm + np.diag(diag)
        z = S.precond_grid(g, h, **kw)
        back = np.einsum("...ij,...j->...i", A, np.moveaxis(z, 1, -1))
        assert rel(back, np.moveaxis(g, 1, -1)) < 1e-5


def test_pcg_converges_to_cg():
    """Block-Jacobi PCG reaches the exact (CG to 1e-12) solution, with the
    energy norm of the error decreasing at every iteration."""
    rng = np.random.default_rng(12)
    shape = (10, 9, 8)
    g = rng.standard_normal((1, 3) + shape)
    h = rng.uniform(0.5, 2, (1, 6) + shape)
    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
    kw = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2), bound="dct2")
    xs, _ = S.cg_grid(h, g, tol=1e-12, **kw)
    prev = np.inf
    for it in (1, 2, 4, 8, 16):
        x, _ = S.pcg_grid(h, g, nb_iter=it, **kw)
        e = np.sum((x - xs) * S.regulariser_grid(x - xs, h, **kw))
        assert e < prev
        prev = e
    x, _ = S.pcg_grid(h, g, nb_iter=400, **kw)
    assert rel(x, xs) < 1e-6


@pytest.mark.parametrize("pen", [dict(absolute=0.3), dict(membrane=0.7, absolute=0.01),
                                 dict(bending=0.2, membrane=0.01, absolute=1e-3)])
def test_plane_wave_eigenvalues_dft(pen):
    """With circulant bounds each component is a plane wave eigenvector of the
    absolute/membrane/bending operator: L e_k = lam(k) e_k,
    lam = v_c^2 (a + m S + b S^2) with S = sum_d (2 - 2 cos t_d) / v_d^2."""
    shape = (8, 6, 10)
    vs = np.array([1.0, 1.5, 0.7])
    k = (3, 1, 2)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    th = [2 * np.pi * k[d] / shape[d] for d in range(3)]
    wave = np.cos(sum(th[d] * grids[d] for d in range(3)))
    Ssym = sum((2 - 2 * np.cos(th[d])) / vs[d] ** 2 for d in range(3))
    a, m, b = pen.get("absolute", 0), pen.get("membrane", 0), pen.get("bending", 0)
    for c in range(3):
        x = np.zeros((1, 3) + shape)
        x[0, c] = wave
        out = S.regulariser_grid(x, None, voxel_size=vs, bound="dft", **pen)
        lam = vs[c] ** 2 * (a + m * Ssym + b * Ssym ** 2)
        assert np.abs(out[0, c] - lam * wave).max() < 1e-10
        others = [d for d in range(3) if d != c]
        assert np.abs(out[0, others]).max() < 1e-12


def test_lame_symbol_dft():
    """Linear elasticity under circulant bounds: for a plane wave u = a cos(k.r),
    L u = (mu |s|^2 a + (mu + lam) s (s.a)) cos(k.r) with the centred-difference
    symbol of the discrete operator (unit voxels)."""
    shape = (8, 8, 8)
    k = (1, 2, 3)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    th = np.array([2 * np.pi * k[d] / shape[d] for d in range(3)])
    phase = sum(th[d] * grids[d] for d in range(3))
    mu, lam = 0.3, 0.2
    for amp in (np.array([1.0, 0, 0]), np.array([0, 1.0, 0.5])):
        x = np.stack([a * np.cos(phase) for a in amp])[None]
        out = S.regulariser_grid(x, None, lame=(mu, lam), bound="dft")
        # second-difference symbols: D_dd = 2 - 2 cos t_d, mixed D_de = sin t_d sin t_e
        Dd = 2 - 2 * np.cos(th)
        M = np.zeros((3, 3))
        for d in range(3):
            M[d, d] = mu * (Dd.sum() + Dd[d]) + lam * Dd[d]
            for e in range(3):
                if e != d:
                    M[d, e] = (mu + lam) * np.sin(th[d]) * np.sin(th[e])
        ref = (M @ amp)[:, None, None, None] * np.cos(phase)[None]
        assert np.abs(out[0] - ref).max() < 1e-10


# With reflecting bounds the reference's Lame cross terms (mixed differences of
# the other components) are not symmetric at the border, and replicate
# (clamping) and dct1 (mirror about the edge voxel centre) are not symmetric
# extensions under the plain inner product; the operator is symmetric for
# circulant bounds, and for reflecting bounds without the Lame term.
@pytest.mark.parametrize("bound,lame", [("dft", (0.05, 0.2)), ("dct2", 0), ("dst2", 0)])
def test_operator_symmetric_positive(bound, lame):
    rng = np.random.default_rng(2)
    shape = (6, 7, 5)
    h = rng.uniform(0.5, 2, (1, 6) + shape)
    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
    kw = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=lame, bound=bound)
    u = rng.standard_normal((1, 3) + shape)
    v = rng.standard_normal((1, 3) + shape)
    Au = S.regulariser_grid(u, h, **kw)
    Av = S.regulariser_grid(v, h, **kw)
    assert abs(np.sum(Au * v) - np.sum(u * Av)) < 1e-10 * abs(np.sum(Au * v))
    assert np.sum(Au * u) > 0


def test_transfer_constants():
    """Prolongation (quadratic, partition of unity) preserves constants for any
    shape; restriction does for exact factors of 2 (its 4 linear weights
    w(d/s)/s sum to 1 when s = 2)."""
    for shape in [(8, 6, 10), (7, 5, 9)]:
        c = np.full((1, 3) + shape, 2.5)
        cs = tuple(1 + (s - 1) // 2 for s in shape)
        if all(s % 2 == 0 for s in shape):
            assert np.abs(S.restrict(c, cs, "dct2") - 2.5).max() < 1e-12
        cc = np.full((1, 3) + cs, -1.5)
        assert np.abs(S.prolong(cc, shape, "dct2") + 1.5).max() < 1e-12


def test_relax_and_fmg_converge_to_cg():
    rng = np.random.default_rng(3)
    shape = (16, 12, 14)
    g = rng.standard_normal((1, 3) + shape)
    h = rng.uniform(0.5, 2, (1, 6) + shape)
    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
    kw = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2), bound="dft")
    xs, it = S.cg_grid(h, g, tol=1e-12, **kw)
    r = g - S.regulariser_grid(xs, h, **kw)
    assert np.linalg.norm(r) < 1e-10 * np.linalg.norm(g)
    # Gauss-Seidel: residual decreases monotonically in energy and converges
    x = np.zeros_like(g)
    prev = np.inf
    for _ in range(5):
        x = S.relax_grid(g, h, x, nb_iter=4, **kw)
        e = np.sum((x - xs) * S.regulariser_grid(x - xs, h, **kw))
        assert e < prev
        prev = e
    # FMG with the intended per-channel scaling: more V-cycles, smaller
    # residuals (the warm start only seeds the coarse levels: the finest
    # solution is replaced by the prolonged coarse one, fmg_common.cpp:199)
    res = []
    for cycles in (1, 2, 4):
        x = S.fmg_grid(h, g, nb_cycles=cycles, nb_iter=4, **kw)
        res.append(np.linalg.norm(g - S.regulariser_grid(x, h, **kw)) / np.linalg.norm(g))
    assert res[0] > res[1] > res[2] and res[2] < 1e-4


BAND_PEN = {3: dict(absolute=0.01, membrane=0.1, bending=0.3, lame=(0.05, 0.2)),
            2: dict(absolute=0.01, membrane=0.1, lame=(0.1, 0.3)),
            0: dict(absolute=0.01, membrane=0.2)}


def _coupling_pairs(shape, bound, pen):
    """Voxel pairs (i, j), i != j, that the operator (H + L) couples: the
    regulariser applied to every unit impulse (one per voxel and component,
    as a batch), with random voxel sizes so that no weight vanishes."""
    X, Y, Z = shape
    V = X * Y * Z
    e = np.zeros((3 * V, 3, V))
    for d in range(3):
        e[np.arange(V) + d * V, d, np.arange(V)] = 1.0
    e = e.reshape(3 * V, 3, X, Y, Z)
    out = S.regulariser_grid(e, None, voxel_size=(0.9, 1.1, 1.3), bound=bound, **pen)
    reach = np.abs(out).reshape(3, V, 3, V).max((0, 2)) > 0      # [source voxel, target voxel]
    np.fill_diagonal(reach, False)
    return np.argwhere(reach)


def _linear_colours(shape, m, abc):
    x, y, z = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return ((abc[0] * x + abc[1] * y + abc[2] * z) % m).ravel()


@pytest.mark.parametrize("bound", ["dft", "dct2", "dst1", "replicate"])
@pytest.mark.parametrize("band", [3, 2, 0])
def test_relax_colour_classes_are_independent(bound, band):
    """The product's colour classes never hold two voxels that the stencil
    couples (checked on the operator's actual coupling graph), so a sweep
    over a class is order-free: visiting each class in reverse gives the
    same iterate, bit for bit.  Both linear colourings and per-axis classes
    (shapes without a linear colouring) are covered."""
    rng = np.random.default_rng(band)
    pen = BAND_PEN[band]
    kinds = set()
    shapes = [(8, 8, 8), (8, 6, 10), (5, 6, 4), (7, 5, 3)]
    if bound == "dft":
        shapes.append((10, 12, 10))     # bending: 15 classes, z coefficient 3
    for shape in shapes:
        kind, m, abc = S.relax_plan(shape, bound, band)
        kinds.add(kind)
        if kind == "linear" and (bound == "dft" or shape[0] * shape[1] * shape[2] <= 200):
            col = _linear_colours(shape, m, abc)
            pairs = _coupling_pairs(shape, bound, pen)
            assert len(pairs) > 0
            assert not (col[pairs[:, 0]] == col[pairs[:, 1]]).any(), (shape, m, abc)
        g = rng.standard_normal((2, 3) + shape)
        h = np.concatenate([rng.uniform(0.5, 2, (2, 3) + shape), rng.uniform(-0.2, 0.2, (2, 3) + shape)], 1)
        x0 = rng.standard_normal((2, 3) + shape)
        a = S.relax_grid(g, h, x0, bound=bound, nb_iter=2, product_order=True, **pen)
        b = S.relax_grid(g, h, x0, bound=bound, nb_iter=2, product_order=True, reverse_within_classes=True, **pen)
        assert np.array_equal(a, b), (shape, kind, m)
        assert not np.array_equal(a, x0)
    if bound == "dft" and band != 0:
        assert kinds == {"linear", "axes"}


def test_relax_colouring_check_detects_conflicts():
    """Planted mistake: the 8-colouring (2x + 3y + z) mod 8 of a 16^3 dft
    lattice used on a 10^3 dft lattice (8 does not divide 10) puts coupled
    voxels across the wrap into one class; the coupling-graph check sees it,
    and so does a sweep in reverse class order."""
    shape = (10, 10, 10)
    assert S.relax_plan((16, 16, 16), "dft", 3)[:2] == ("linear", 8)
    col = _linear_colours(shape, 8, (2, 3, 1))
    pairs = _coupling_pairs(shape, "dft", BAND_PEN[3])
    assert (col[pairs[:, 0]] == col[pairs[:, 1]]).any()
    # and the product plan for 10^3 differs from that colouring
    assert S.relax_plan(shape, "dft", 3) != ("linear", 8, (2, 3, 1))


def test_relax_plan_sizes():
    """Class counts of the product's plans on the fmg_160 pyramid (dft,
    bending): 8 classes on 160x192x160 .. 40x48x40, 15 below; 7 with
    reflecting bounds; red/black for membrane only."""
    for shape in [(160, 192, 160), (80, 96, 80), (40, 48, 40)]:
        assert S.relax_plan(shape, "dft", 3) == ("linear", 8, (2, 3, 1))
        assert S.relax_plan(shape, "dct2", 3) == ("linear", 7, (2, 3, 1))
        assert S.relax_plan(shape, "dft", 0) == ("linear", 2, (1, 1, 1))
    for shape in [(20, 24, 20), (10, 12, 10)]:
        assert S.relax_plan(shape, "dft", 3)[:2] == ("linear", 15)
    # no linear colouring of an odd circulant lattice with <= 60 classes
    assert S.relax_plan((7, 5, 9), "dft", 3)[0] == "axes"
