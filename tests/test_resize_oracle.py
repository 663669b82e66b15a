"""The resize oracle (oracle/resize.py) against the reference's own compiled
resize (nitorch/_C/src/resize.cpp, CPU build in oracle/_ref): every order
0-7 (isotropic and mixed), every alignment mode, enlarging, shrinking and
anisotropic factors, the adjoint, `normalize`, and given outputs (added
into or overwritten).  The reference allocates 5-D outputs whatever the
dimension, so its 1-D / 2-D calls fail without a given output: the 3-D
cases are the pinned ones.  Two identities that do not depend on the
reference close the loop: the resize and its adjoint are transposes, and a
center-aligned linear resize keeps the corner samples."""
import numpy as np
import pytest
import torch

from oracle import ref as R
from oracle import resize as RS

NAMES = {0: "edge", 1: "center", 2: "first", 3: "last"}


def _ref_resize(m, x, factor, bound, order, mode, out, adjoint, normalize):
    o = torch.from_numpy(out.copy()) if out is not None else torch.Tensor()
    return m.resize(torch.from_numpy(x.copy()), o, list(factor), R.bounds(bound), R.orders(order),
                    [getattr(m.GridAlignType, NAMES[md]) for md in mode], adjoint, normalize).numpy()


@pytest.mark.skipif(not R.available(), reason="reference CPU build (oracle/_ref) not built")
@pytest.mark.parametrize("order", [[0], [1], [2], [3], [5], [7], [1, 3, 0], [2, 2, 4]])
def test_matches_reference(order):
    m = R.load()
    rng = np.random.default_rng(sum(order))
    for mode in ([0], [1], [2], [3], [0, 1, 2]):
        for f in ([2.0], [0.5], [1.5], [0.7, 1.3, 2.0]):
            for adjoint in (False, True):
                for normalize in (False, True):
                    for given in (False, True):
                        x = rng.standard_normal((2, 3, 7, 6, 5))
                        bound = ["dct2", "dft", "replicate"]
                        shp = RS.out_shape(x.shape[2:], f, adjoint)
                        out = rng.standard_normal((2, 3) + tuple(shp)) if given else None
                        ref = _ref_resize(m, x, f, bound, order, mode, out, adjoint, normalize)
                        got = RS.resize(x, f, bound, order, mode, None, out, adjoint, normalize)
                        assert got.shape == ref.shape
                        assert np.allclose(got, ref, rtol=1e-11, atol=1e-11), (mode, f, adjoint, normalize, given)


@pytest.mark.skipif(not R.available(), reason="reference CPU build (oracle/_ref) not built")
def test_other_bounds_match_reference():
    m = R.load()
    rng = np.random.default_rng(9)
    for bound in (["zero"], ["dct1"], ["dst1"], ["dst2"], ["dft", "zero", "dct1"]):
        x = rng.standard_normal((1, 2, 6, 7, 5))
        for adjoint in (False, True):
            ref = _ref_resize(m, x, [1.7], bound, [3], [0], None, adjoint, False)
            got = RS.resize(x, [1.7], bound, [3], [0], None, None, adjoint, False)
            assert np.allclose(got, ref, rtol=1e-11, atol=1e-11), (bound, adjoint)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_adjoint_is_transpose(dim):
    rng = np.random.default_rng(dim)
    shp = (5, 7, 6)[:dim]
    x = rng.standard_normal((2, 2) + shp)
    for order in (1, 3, 4):
        for mode in (0, 1, 2):
            y = RS.resize(x, [1.6], "dct2", order, mode)
            g = rng.standard_normal(y.shape)
            # the adjoint inverts its factor (resize_backward passes 1 / factor)
            back = RS.resize(g, [1 / 1.6], "dct2", order, mode, shape=shp, adjoint=True)
            assert back.shape == x.shape
            assert abs((y * g).sum() - (x * back).sum()) <= 1e-10 * abs((y * g).sum())


def test_center_linear_keeps_corners():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, 1, 6, 5, 7))
    y = RS.resize(x, [2.5], "dct2", 1, RS.CENTER)
    for idx in [(0, 0, 0), (-1, -1, -1), (0, -1, 0), (-1, 0, -1)]:
        assert np.isclose(y[(0, 0) + idx], x[(0, 0) + idx], atol=1e-12)
