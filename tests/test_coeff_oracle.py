"""The prefilter oracle (oracle/coeff.py) pinned against what fixes it
independently of the reference's text:

* scipy.ndimage.spline_filter1d (orders 2-5, published Thevenaz/Unser
  recursions): mirror == dct1 and grid-wrap == dft at every length;
  reflect == dct2 on lines long enough for the reference's rewritten
  initial value to agree (on short lines the two differ by O(z^(2n-1)),
  which is bounded here);
* the interpolation property for orders 2-7: the oracle's own pull of the
  coefficients at the sample positions returns the samples (dct1, dft, and
  dct2 on long lines) -- this pins the order 6 and 7 poles scipy lacks;
* a pole perturbed in the oracle breaks that property (the check is
  sensitive), and orders 0/1 and length-1 lines are the identity.
"""
import math

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import coeff as C
from oracle import pushpull as PP

MODES = {"dct1": "mirror", "dft": "grid-wrap", "dct2": "reflect"}


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_matches_scipy(order):
    rng = np.random.default_rng(order)
    for bound, mode in MODES.items():
        lengths = [16, 31, 64, 100] if bound == "dct2" else [2, 3, 4, 5, 7, 12, 31, 64, 100]
        for n in lengths:
            x = rng.standard_normal((3, n))
            ref = ndi.spline_filter1d(x, order, axis=-1, mode=mode)
            out = C.filter1d(x, bound, order)
            assert np.abs(out - ref).max() <= 1e-11 * np.abs(ref).max(), (bound, n)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_dct2_short_lines_differ_by_tail(order):
    """The reference's dct2 initial value (coeff.py:141-176) is scipy's with
    terms of order z^(2n-1) rearranged: the gap is bounded by that order."""
    rng = np.random.default_rng(10 + order)
    zmax = max(abs(z) for z in C.poles(order))
    for n in (3, 4, 5, 8, 12):
        x = rng.standard_normal((4, n))
        gap = np.abs(C.filter1d(x, "dct2", order) - ndi.spline_filter1d(x, order, axis=-1, mode="reflect")).max()
        assert gap <= 50 * zmax ** (2 * n - 1) * np.abs(x).max() * C.gain(C.poles(order)), n


def _identity(shape):
    return np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij"), -1)[None]


def _interp_error(x, bound, order):
    c = C.spline_coeff_nd(x, bound, order, dim=3)
    back = PP.grid_pull(c, _identity(x.shape[2:]), bound, order, 0)
    return np.abs(back - x).max() / np.abs(x).max()


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 7])
def test_interpolation_property(order):
    rng = np.random.default_rng(20 + order)
    for bound, shape in [("dct1", (7, 5, 9)), ("dft", (6, 9, 4)), ("dct2", (24, 26, 23)),
                         ("dct1", (2, 3, 60)), ("dft", (60, 2, 3))]:
        x = rng.standard_normal((1, 2) + shape)
        assert _interp_error(x, bound, order) <= 1e-9, (bound, shape)


def test_perturbed_pole_is_caught(monkeypatch):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 1, 9, 8, 7))
    assert _interp_error(x, "dct1", 7) <= 1e-9
    good = C.poles

    def bad(order):
        z = list(good(order))
        z[-1] *= 1.0 + 1e-6
        return z
    monkeypatch.setattr(C, "poles", bad)
    assert _interp_error(x, "dct1", 7) > 1e-8


def test_identity_cases_and_rules():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 5, 1, 6))
    for order in (0, 1):
        assert np.array_equal(C.spline_coeff_nd(x, "dct2", order, dim=3), x)
    # a length-1 axis is left unchanged, the others are filtered
    y = C.spline_coeff_nd(x, "dct2", 3, dim=3)
    assert np.allclose(y, C.filter1d(C.filter1d(x, "dct2", 3, axis=1), "dct2", 3, axis=3))
    # zero uses the dct1 rules, replicate the dct2 ones; dst bounds are refused
    assert np.array_equal(C.filter1d(x, "zero", 3, axis=3), C.filter1d(x, "dct1", 3, axis=3))
    assert np.array_equal(C.filter1d(x, "replicate", 3, axis=3), C.filter1d(x, "dct2", 3, axis=3))
    with pytest.raises(NotImplementedError):
        C.filter1d(x, "dst2", 3, axis=3)
    # gain and horizon of the reference's definitions
    z = math.sqrt(3.) - 2.
    assert C.gain([z]) == pytest.approx(6.0, rel=1e-14)
    assert C.horizon(z) == math.ceil(-30. / math.log(abs(z)))
