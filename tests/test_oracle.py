"""CPU tests of the oracle (plain-C restatement of the reference's push/pull).

The oracle is pinned by
  * the golden vectors written by the reference's own compiled CPU backend
    (tests/golden/pushpull_golden.npz, tests/golden/make_golden.py);
  * closed forms and invariants that a plausible mistake would break:
    partition of unity, derivative identities checked by finite differences,
    boundary tables checked against numpy.pad modes and the reference's
    documented tables, identity-grid pulls, torch.grid_sample for orders 0/1,
    exact pull/push adjointness, count mass conservation;
  * (when oracle/_ref is built) random cases against the reference itself.
"""
import os

import numpy as np
import pytest
import torch

from oracle import pushpull as O
from oracle import ref as R

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "pushpull_golden.npz")
BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]


def rel(a, b):
    den = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(np.asarray(a, np.float64) - b).max()) / den


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def _cases(g):
    keys = sorted({k.split("_")[0] for k in g.files})
    for k in keys:
        yield k, {f.split("_", 1)[1]: g[f] for f in g.files if f.startswith(k + "_")}


def test_golden_pushpull(golden):
    """Oracle == reference outputs on every golden case and entry point."""
    n = 0
    for key, c in _cases(golden):
        dim, order, b, extrap, f64 = [int(v) for v in c["meta"]]
        bound = BOUNDS[b]
        tol = 1e-12 if f64 else 2e-6
        inp, grid = c["inp"], c["grid"]
        shape = inp.shape[2:]
        checks = {
            "pull": lambda: O.grid_pull(inp, grid, bound, order, extrap),
            "push": lambda: O.grid_push(c["srcpush"], grid, shape, bound, order, extrap),
            "count": lambda: O.grid_count(grid, shape, bound, order, extrap),
            "pbw_in": lambda: O.grid_pull_backward(c["gpull"], inp, grid, bound, order, extrap, need_grid=False)[0],
            "ubw_in": lambda: O.grid_push_backward(inp, c["srcpush"], grid, bound, order, extrap, need_grid=False)[0],
        }
        if "sgrad" in c:
            checks["sgrad"] = lambda: O.grid_grad(inp, grid, bound, order, extrap)
        if "pbw_grid" in c:
            checks["pbw_grid"] = lambda: O.grid_pull_backward(c["gpull"], inp, grid, bound, order, extrap)[1]
            checks["ubw_grid"] = lambda: O.grid_push_backward(inp, c["srcpush"], grid, bound, order, extrap)[1]
            checks["cbw_grid"] = lambda: O.grid_count_backward(c["gcount"], grid, bound, order, extrap)
        for name, fn in checks.items():
            e = rel(fn(), c[name])
            assert e <= tol, (key, name, dim, order, bound, extrap, f64, e)
            n += 1
    assert n > 2000


def test_order0_bit_exact_against_golden(golden):
    """Nearest-neighbour pulls copy values: bit-exact with the reference."""
    for key, c in _cases(golden):
        dim, order, b, extrap, f64 = [int(v) for v in c["meta"]]
        if order != 0:
            continue
        out = O.grid_pull(c["inp"], c["grid"], BOUNDS[b], 0, extrap)
        assert np.array_equal(out, c["pull"]), key


@pytest.mark.parametrize("order", range(8))
def test_spline_partition_of_unity(order):
    for x in np.linspace(-0.49, 0.49, 37):
        # taps lo..lo+order around x
        lo = int(np.round(x)) if order == 0 else int(np.floor(x - (order - 1) / 2))
        s = sum(O.spline_weight(order, x - (lo + k)) for k in range(order + 1))
        assert abs(s - 1) < 1e-12
        if order >= 1:
            sg = sum(O.spline_grad(order, x - (lo + k)) for k in range(order + 1))
            assert abs(sg) < 1e-12
        if order >= 2:
            sh = sum(O.spline_hess(order, x - (lo + k)) for k in range(order + 1))
            assert abs(sh) < 1e-10


@pytest.mark.parametrize("order", range(1, 8))
def test_spline_derivatives_match_finite_differences(order):
    h = 1e-6
    for x in np.linspace(-(order + 1) / 2 + 0.013, (order + 1) / 2 - 0.013, 60):
        fd = (O.spline_weight(order, x + h) - O.spline_weight(order, x - h)) / (2 * h)
        assert abs(fd - O.spline_grad(order, x)) < 1e-6, (order, x)
        if order >= 3:
            fd2 = (O.spline_grad(order, x + h) - O.spline_grad(order, x - h)) / (2 * h)
            assert abs(fd2 - O.spline_hess(order, x)) < 1e-5, (order, x)


def test_spline_closed_forms():
    # linear and cubic closed forms
    for x in np.linspace(-1.9, 1.9, 39):
        assert abs(O.spline_weight(1, x) - max(0, 1 - abs(x))) < 1e-14
        a = abs(x)
        c = (2 / 3 - a * a + a ** 3 / 2) if a < 1 else ((2 - a) ** 3 / 6 if a < 2 else 0)
        assert abs(O.spline_weight(3, x) - c) < 1e-14
    # nearest: half-open support
    assert O.spline_weight(0, 0.49) == 1 and O.spline_weight(0, 0.5) == 0


@pytest.mark.parametrize("order", range(1, 8))
def test_spline_last_piece_relative_accuracy(order):
    """On the last piece the B-spline is (h - |x|)^n / n! (h = (n + 1) / 2);
    near the end of the support the weight is tiny and must keep its
    relative accuracy (a truncated-power or Horner form cancels there)."""
    import math
    h = (order + 1) / 2
    for eps in (1e-7, 1e-5, 1e-3, 0.1, 0.7):
        for x in (h - eps, -(h - eps)):
            s = h - abs(x)   # exact (Sterbenz)
            want = s ** order / math.factorial(order)
            assert abs(O.spline_weight(order, x) / want - 1) < 1e-13, (order, x)


@pytest.mark.parametrize("bound,mode", [("dct2", "symmetric"), ("dct1", "reflect"),
                                        ("dft", "wrap"), ("replicate", "edge")])
def test_bound_index_matches_numpy_pad(bound, mode):
    for n in (1, 2, 3, 5):
        if n == 1 and mode == "reflect":
            continue
        pad = 2 * n + 3
        ref = np.pad(np.arange(n), pad, mode=mode)
        for i in range(-pad, n + pad):
            assert O.bound_index(bound, i, n) == ref[i + pad], (bound, n, i)
            assert O.bound_sign(bound, i, n) == 1


def test_bound_tables_documented():
    """The tables of nitorch/_C/spatial.cpp:48-62 for a b c d (n=4)."""
    n = 4
    abcd = np.array([1, 2, 3, 4])

    def seq(b):
        return [O.bound_sign(b, i, n) * (abcd[O.bound_index(b, i, n)] if O.bound_sign(b, i, n) else 1)
                for i in range(-3, n + 3)]
    a, b_, c, d = 1, 2, 3, 4
    assert seq("dst1") == [-b_, -a, 0, a, b_, c, d, 0, -d, -c]
    assert seq("dst2") == [-c, -b_, -a, a, b_, c, d, -d, -c, -b_]
    assert seq("dct1") == [d, c, b_, a, b_, c, d, c, b_, a]
    assert seq("dct2") == [c, b_, a, a, b_, c, d, d, c, b_]
    assert seq("dft") == [b_, c, d, a, b_, c, d, a, b_, c]
    z = [O.bound_sign("zero", i, n) for i in range(-3, n + 3)]
    assert z == [0, 0, 0, 1, 1, 1, 1, 0, 0, 0]


def _identity(shape):
    return np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij"), -1)[None]


@pytest.mark.parametrize("bound", BOUNDS)
def test_pull_identity(bound):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 2, 6, 5, 7))
    g = _identity((6, 5, 7))
    for order in (0, 1):
        assert np.abs(O.grid_pull(x, g, bound, order) - x).max() < 1e-14
    # cubic: separable FIR [1/6, 2/3, 1/6] with the boundary condition
    if bound in ("dct2", "dft", "replicate"):
        mode = {"dct2": "symmetric", "dft": "wrap", "replicate": "edge"}[bound]
        ref = x
        for ax in (2, 3, 4):
            p = np.pad(ref, [(0, 0)] * ax + [(1, 1)] + [(0, 0)] * (4 - ax), mode=mode)
            sl = lambda a, b: tuple([slice(None)] * ax + [slice(a, b)])
            n = ref.shape[ax]
            ref = p[sl(0, n)] / 6 + p[sl(1, n + 1)] * 2 / 3 + p[sl(2, n + 2)] / 6
        assert np.abs(O.grid_pull(x, g, bound, 3) - ref).max() < 1e-13


# torch's "reflection" with align_corners=True mirrors about the centres of the
# edge voxels, i.e. nitorch's dct1 (the reference's documentation table in
# nitorch/_C/spatial.cpp:30-36 pairs it with dct2; the arithmetic says dct1).
@pytest.mark.parametrize("bound,pad", [("zero", "zeros"), ("replicate", "border"), ("dct1", "reflection")])
@pytest.mark.parametrize("order", [0, 1])
def test_pull_matches_grid_sample(bound, pad, order):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 2, 6, 7, 5))
    g = _identity((4, 5, 6)) * 0.9 + rng.uniform(-0.6, 0.6, (1, 4, 5, 6, 3))
    if order == 0:   # keep clear of .5 ties (grid_sample rounds half to even)
        frac = g - np.floor(g)
        g = np.where(np.abs(frac - 0.5) < 0.05, g + 0.1, g)
    out = O.grid_pull(x, g, bound, order, extrapolate=1)
    sz = np.array(x.shape[2:])
    # [N, C, X, Y, Z] is torch's [N, C, D, H, W]; torch's grid holds (W, H, D)
    # coordinates normalised to [-1, 1] (align_corners=True)
    gn = 2 * g / (sz - 1) - 1
    gt = torch.from_numpy(gn[..., ::-1].copy())
    ref = torch.nn.functional.grid_sample(torch.from_numpy(x), gt, mode="nearest" if order == 0 else "bilinear",
                                          padding_mode=pad, align_corners=True).numpy()
    assert np.abs(out - ref).max() < 1e-12


@pytest.mark.parametrize("order", range(8))
@pytest.mark.parametrize("bound", BOUNDS)
def test_pull_push_adjoint(order, bound):
    rng = np.random.default_rng(order * 7 + BOUNDS.index(bound))
    x = rng.standard_normal((2, 3, 5, 6, 4))
    g = _identity((4, 3, 5)) + 1.5 * rng.standard_normal((2, 4, 3, 5, 3))
    y = rng.standard_normal((2, 3, 4, 3, 5))
    for extrap in (0, 1):
        lhs = np.sum(O.grid_pull(x, g, bound, order, extrap) * y)
        rhs = np.sum(x * O.grid_push(y, g, x.shape[2:], bound, order, extrap))
        assert abs(lhs - rhs) <= 1e-12 * max(1, abs(lhs))
        # count = push of ones
        c = O.grid_count(g, x.shape[2:], bound, order, extrap)
        p = O.grid_push(np.ones((2, 1) + g.shape[1:-1]), g, x.shape[2:], bound, order, extrap)
        assert np.abs(c - p).max() < 1e-12


@pytest.mark.parametrize("bound", ["dct2", "dct1", "dft", "replicate"])
def test_count_mass(bound):
    rng = np.random.default_rng(3)
    g = _identity((6, 5, 4)) + rng.standard_normal((1, 6, 5, 4, 3))
    sz = (6, 5, 4)
    for order in (1, 2, 3, 5):
        c = O.grid_count(g, sz, bound, order, 0)
        tol = 0.05
        inb = np.all((g >= -tol) & (g < np.array(sz) - 1 + tol), -1).sum()
        assert abs(c.sum() - inb) < 1e-10


def test_backward_matches_finite_differences():
    """Grid gradient of <pull(x, g), y> against central differences (FP64)."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, 2, 6, 5, 7))
    g = _identity((3, 4, 3)) + 1.0 + 0.4 * rng.standard_normal((1, 3, 4, 3, 3))
    y = rng.standard_normal((1, 2, 3, 4, 3))
    for order in (2, 3, 5):
        for bound in ("dct2", "dft"):
            _, gg = O.grid_pull_backward(y, x, g, bound, order, 1)
            d = rng.standard_normal(g.shape)
            h = 1e-5
            fp = np.sum(O.grid_pull(x, g + h * d, bound, order, 1) * y)
            fm = np.sum(O.grid_pull(x, g - h * d, bound, order, 1) * y)
            assert abs((fp - fm) / (2 * h) - np.sum(gg * d)) < 1e-6 * max(1, abs(np.sum(gg * d)))


def test_grad_backward_hessian_finite_differences():
    """grid_grad_backward's grid part (2nd derivatives, all orders) against FD."""
    rng = np.random.default_rng(6)
    x = rng.standard_normal((1, 2, 7, 6, 7))
    g = _identity((3, 3, 3)) + 1.5 + 0.3 * rng.standard_normal((1, 3, 3, 3, 3))
    y = rng.standard_normal((1, 2, 3, 3, 3, 3))
    for order in (3, 4, 5, 6, 7):
        _, gg = O.grid_grad_backward(y, x, g, "dft", order, 1)
        d = rng.standard_normal(g.shape)
        h = 1e-5
        fp = np.sum(O.grid_grad(x, g + h * d, "dft", order, 1) * y)
        fm = np.sum(O.grid_grad(x, g - h * d, "dft", order, 1) * y)
        assert abs((fp - fm) / (2 * h) - np.sum(gg * d)) < 1e-5 * max(1, abs(np.sum(gg * d))), order


@pytest.mark.skipif(not R.available(), reason="reference backend not built (oracle/Makefile.ref)")
def test_oracle_vs_reference_random():
    m = R.load()
    rng = np.random.default_rng(11)
    for trial in range(40):
        dim = int(rng.integers(2, 4))
        order = int(rng.integers(0, 8))
        bound = BOUNDS[int(rng.integers(0, 7))]
        sp = tuple(int(v) for v in rng.integers(3, 7, dim))
        x = rng.standard_normal((2, 2) + sp)
        g = _identity(sp) + 1.2 * rng.standard_normal((2,) + sp + (dim,))
        r = m.grid_pull(torch.from_numpy(x), torch.from_numpy(g), R.bounds([bound]), R.orders([order]), 0).numpy()
        assert rel(O.grid_pull(x, g, bound, order, 0), r) < 1e-12
        y = rng.standard_normal(x.shape)
        r = m.grid_push(torch.from_numpy(y), torch.from_numpy(g), list(sp), R.bounds([bound]),
                        R.orders([order]), 0).numpy()
        assert rel(O.grid_push(y, g, sp, bound, order, 0), r) < 1e-12


GOLDEN_EXTRA = os.path.join(HERE, "golden", "pushpull_golden_extra.npz")


def test_golden_extrapolate2_and_grad_backward_input():
    """Oracle == reference on the extrapolate=2 band (tolerance 0.55) for every
    entry point, and on grid_grad_backward's input branch at extrapolate
    0 / 1 / 2 (tests/golden/make_golden_extra.py)."""
    g = np.load(GOLDEN_EXTRA)
    n = n_band = 0
    for key, c in _cases(g):
        dim, order, b, extrap, f64 = [int(v) for v in c["meta"]]
        assert extrap == 2
        bound = BOUNDS[b]
        tol = 1e-12 if f64 else 2e-6
        inp, grid = c["inp"], c["grid"]
        shape = inp.shape[2:]
        # the grid fills the band between the two tolerances on some axis
        sz = np.array(shape, np.float64)
        band = np.any(((grid >= -0.55) & (grid < -0.05)) | ((grid >= sz - 1 + 0.05) & (grid < sz - 1 + 0.55)), -1)
        n_band += int(band.sum())
        checks = {
            "pull": lambda: O.grid_pull(inp, grid, bound, order, 2),
            "push": lambda: O.grid_push(c["srcpush"], grid, shape, bound, order, 2),
            "count": lambda: O.grid_count(grid, shape, bound, order, 2),
            "pbw_in": lambda: O.grid_pull_backward(c["gpull"], inp, grid, bound, order, 2, need_grid=False)[0],
            "ubw_in": lambda: O.grid_push_backward(inp, c["srcpush"], grid, bound, order, 2, need_grid=False)[0],
        }
        if "sgrad" in c:
            checks["sgrad"] = lambda: O.grid_grad(inp, grid, bound, order, 2)
        if "pbw_grid" in c:
            checks["pbw_grid"] = lambda: O.grid_pull_backward(c["gpull"], inp, grid, bound, order, 2)[1]
            checks["ubw_grid"] = lambda: O.grid_push_backward(inp, c["srcpush"], grid, bound, order, 2)[1]
            checks["cbw_grid"] = lambda: O.grid_count_backward(c["gcount"], grid, bound, order, 2)
        for ex in (0, 1, 2):
            if f"gbw_in{ex}" in c:
                checks[f"gbw_in{ex}"] = (lambda ex=ex: O.grid_grad_backward(c["ggrad"], inp, grid, bound, order, ex,
                                                                          need_grid=False)[0])
        for name, fn in checks.items():
            ref = c[name]
            e = rel(fn(), ref) if np.abs(ref).max() > 0 else float(np.abs(fn()).max())
            assert e <= tol, (key, name, dim, order, bound, f64, e)
            n += 1
    assert n > 2000 and n_band > 500, (n, n_band)


def test_extrapolate2_tolerance_closed_form():
    """The widened tolerance on its own: a sample 0.3 voxel outside the first
    or last voxel is out of bounds at extrapolate=0 (pull gives 0) and in
    bounds at extrapolate=2, where linear interpolation with replicate
    padding returns the edge value exactly."""
    x = np.arange(1.0, 7.0).reshape(1, 1, 6)
    grid = np.array([-0.3, 5.3, -0.6, 2.5]).reshape(1, 4, 1)
    p0 = O.grid_pull(x, grid, "replicate", 1, 0)[0, 0]
    p2 = O.grid_pull(x, grid, "replicate", 1, 2)[0, 0]
    assert p0[0] == 0 and p0[1] == 0 and p0[3] == 3.5
    assert p2[0] == 1.0 and p2[1] == 6.0 and p2[2] == 0 and p2[3] == 3.5


GOLDEN_EDGE = os.path.join(HERE, "golden", "pushpull_golden_edge.npz")


def test_golden_float_edges_oob_f32():
    """float32 coordinates exactly on the out-of-bound edges (and their float
    neighbours), extrapolate 0 and 2: the FP64 oracle with `oob_f32` (the
    in-bound test in float, as the reference's float build does it) matches
    the reference's float32 outputs, with the same zero pattern of the pull
    (out-of-bound samples give exactly 0); the FP64 test alone disagrees on
    some of these samples (tests/golden/make_golden_edge.py)."""
    g = np.load(GOLDEN_EDGE)
    n = n_diff = 0
    for key, c in _cases(g):
        dim, order, b, extrap = [int(v) for v in c["meta"]]
        bound = BOUNDS[b]
        inp, grid = c["inp"].astype(np.float64), c["grid"].astype(np.float64)
        src = c["srcpush"].astype(np.float64)
        shape = inp.shape[2:]
        pull = O.grid_pull(inp, grid, bound, order, extrap, oob_f32=True)
        assert np.array_equal(pull == 0, c["pull"] == 0), key
        got = {"pull": pull,
               "push": O.grid_push(src, grid, shape, bound, order, extrap, oob_f32=True),
               "count": O.grid_count(grid, shape, bound, order, extrap, oob_f32=True)}
        for name, out in got.items():
            assert rel(out, c[name]) <= 2e-6, (key, name, order, bound, extrap)
            n += 1
        plain = O.grid_pull(inp, grid, bound, order, extrap)
        n_diff += int(((plain == 0) != (pull == 0)).sum())
    assert n == 216 and n_diff > 50, (n, n_diff)


@pytest.mark.parametrize("order", range(8))
@pytest.mark.parametrize("bound", BOUNDS)
def test_grad_grad_backward_input_adjoint(order, bound):
    """<grid_grad(x, g), y> = <x, grid_grad_backward(y, x, g)[input]> in FP64:
    the input branch is the exact adjoint of grid_grad (a dropped tap, a wrong
    sign or a transposed component breaks it), for extrapolate 0 / 1 / 2."""
    rng = np.random.default_rng(100 + order * 7 + BOUNDS.index(bound))
    x = rng.standard_normal((2, 3, 5, 6, 4))
    g = _identity((4, 3, 5)) + 1.5 * rng.standard_normal((2, 4, 3, 5, 3))
    y = rng.standard_normal((2, 3, 4, 3, 5, 3))
    for extrap in (0, 1, 2):
        lhs = np.sum(O.grid_grad(x, g, bound, order, extrap) * y)
        gi, _ = O.grid_grad_backward(y, x, g, bound, order, extrap, need_grid=False)
        rhs = np.sum(x * gi)
        assert abs(lhs - rhs) <= 1e-12 * max(1, np.abs(y).sum()), (extrap, lhs, rhs)
        if order == 0:
            assert np.abs(gi).max() == 0


def test_grad_grad_backward_input_finite_differences():
    """grid_grad is linear in its input, so the input branch equals the
    directional derivative of <grid_grad(x, g), y> along any input direction,
    and grid_grad itself equals central differences of pull along each grid
    axis (orders >= 2, where the spline derivative is continuous)."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal((1, 2, 7, 6, 5))
    g = _identity((3, 3, 3)) + 1.5 + 0.3 * rng.standard_normal((1, 3, 3, 3, 3))
    y = rng.standard_normal((1, 2, 3, 3, 3, 3))
    for order in (2, 3, 5, 7):
        gi, _ = O.grid_grad_backward(y, x, g, "dct2", order, 1, need_grid=False)
        d = rng.standard_normal(x.shape)
        h = 1e-3
        fp = np.sum(O.grid_grad(x + h * d, g, "dct2", order, 1) * y)
        fm = np.sum(O.grid_grad(x - h * d, g, "dct2", order, 1) * y)
        assert abs((fp - fm) / (2 * h) - np.sum(gi * d)) < 1e-9 * max(1, abs(np.sum(gi * d)))
        sg = O.grid_grad(x, g, "dct2", order, 1)
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1e-5
            fd = (O.grid_pull(x, g + e, "dct2", order, 1) - O.grid_pull(x, g - e, "dct2", order, 1)) / 2e-5
            assert np.abs(fd - sg[..., k]).max() < 1e-6 * max(1, np.abs(sg).max()), (order, k)
