"""ORACLE — test infrastructure only.

FP64 numpy restatement of the interpolating B-spline prefilter
(`spline_coeff` / `spline_coeff_nd`), following the reference's in-tree
TorchScript version nitorch/_C/_ts/coeff.py (the compiled backend imports
the same routines from the third-party `interpol.coeff`, torch-interpol,
unpinned in setup.cfg:27 and absent here):

* poles per order (coeff.py:34-57) and gain prod (1 - z)(1 - 1/z) (:60-65);
* causal initial value: dft (:68-91), dct1 (:94-138, truncated sum when the
  horizon ceil(-30 / ln|z|) is shorter than the line), dct2 (:141-176);
* anticausal initial value: dft (:179-201), dct1 (:204-211), dct2 (:214-222);
* bound dispatch (:225-248): zero/dct1 -> dct1 rules, replicate/dct2 -> dct2
  rules, dft -> dft rules, others unsupported;
* filter (:251-275): gain, then per pole a causal and an anticausal pass;
  lines of length 1 are left unchanged;
* spline_coeff_nd (:309-344): the 1-D filter along each of the last `dim`
  axes, first to last; orders 0 and 1 are the identity.
"""
import math

import numpy as np

_DCT1 = ("dct1", "zero")
_DCT2 = ("dct2", "replicate")
_DFT = ("dft",)


def poles(order):
    """Poles of the order-`order` B-spline interpolation filter (coeff.py:34-57)."""
    if order in (0, 1):
        return []
    if order == 2:
        return [math.sqrt(8.) - 3.]
    if order == 3:
        return [math.sqrt(3.) - 2.]
    if order == 4:
        return [math.sqrt(664. - math.sqrt(438976.)) + math.sqrt(304.) - 19.,
                math.sqrt(664. + math.sqrt(438976.)) - math.sqrt(304.) - 19.]
    if order == 5:
        return [math.sqrt(67.5 - math.sqrt(4436.25)) + math.sqrt(26.25) - 6.5,
                math.sqrt(67.5 + math.sqrt(4436.25)) - math.sqrt(26.25) - 6.5]
    if order == 6:
        return [-0.488294589303044755130118038883789062112279161239377608394,
                -0.081679271076237512597937765737059080653379610398148178525368,
                -0.00141415180832581775108724397655859252786416905534669851652709]
    if order == 7:
        return [-0.5352804307964381655424037816816460718339231523426924148812,
                -0.122554615192326690515272264359357343605486549427295558490763,
                -0.0091486948096082769285930216516478534156925639545994482648003]
    raise NotImplementedError(order)


def gain(zs):
    g = 1.0
    for z in zs:
        g *= (1. - z) * (1. - 1. / z)
    return g


def horizon(z):
    return int(math.ceil(-30. / math.log(abs(z))))


# x: [n, ...] (filtered axis first); each returns the new x[0] / x[n-1]
def dft_initial(x, z):
    n = x.shape[0]
    m = min(horizon(z), n)
    out = x[0].copy()
    for k in range(1, m):
        out += z ** k * x[n - k]
    return out / (1 - z ** m)


def dct1_initial(x, z):
    n = x.shape[0]
    m = horizon(z)
    if m < n:
        out = x[0].copy()
        for k in range(1, m):
            out += z ** k * x[k]
        return out
    zn = z ** (n - 1)
    out = x[0] + zn * x[n - 1]
    for k in range(1, n - 1):
        out += (z ** k + zn * zn / z ** k) * x[k]
    return out / (1 - zn * zn)


def dct2_initial(x, z):
    n = x.shape[0]
    zn = z ** n
    last = zn * (1 + 1 / (z + zn * zn))
    out = x[0] + last * x[n - 1]
    for k in range(1, n - 1):
        out += (z ** k + z ** (2 * n - 1 - k)) * x[k]
    return out * (z / (1 - zn * zn)) + x[0]


def dft_final(x, z):
    n = x.shape[0]
    m = min(horizon(z), n)
    out = z * x[n - 1]
    for j in range(0, m - 1):
        out += z ** (j + 2) * x[j]
    return out / (z ** m - 1)


def dct1_final(x, z):
    return (z * x[-2] + x[-1]) * (z / (z * z - 1))


def dct2_final(x, z):
    return x[-1] * (z / (z - 1))


def _rules(bound):
    if bound in _DCT1:
        return dct1_initial, dct1_final
    if bound in _DCT2:
        return dct2_initial, dct2_final
    if bound in _DFT:
        return dft_initial, dft_final
    raise NotImplementedError(f"spline prefilter: bound {bound!r} is not supported")


def filter1d(x, bound, order, axis=-1):
    """Interpolating coefficients along one axis (coeff.py:251-306)."""
    x = np.array(x, dtype=np.float64, copy=True)
    zs = poles(order)
    if not zs or x.shape[axis] == 1:
        return x
    initial, final = _rules(bound)
    x = np.moveaxis(x, axis, 0)
    n = x.shape[0]
    x *= gain(zs)
    for z in zs:
        x[0] = initial(x, z)
        for i in range(1, n):
            x[i] += z * x[i - 1]
        x[n - 1] = final(x, z)
        for i in range(n - 2, -1, -1):
            x[i] = z * (x[i + 1] - x[i])
    return np.moveaxis(x, 0, axis)


def spline_coeff_nd(x, bound, order, dim=None):
    """Along each of the last `dim` axes, first to last (coeff.py:309-344);
    `bound` / `order` lists are padded by repeating their last entry."""
    x = np.asarray(x, np.float64)
    dim = x.ndim if dim is None else dim
    bound = list(bound) if isinstance(bound, (list, tuple)) else [bound]
    order = list(order) if isinstance(order, (list, tuple)) else [order]
    bound = (bound + [bound[-1]] * dim)[:dim]
    order = (order + [order[-1]] * dim)[:dim]
    for d in range(dim):
        x = filter1d(x, bound[d], order[d], axis=x.ndim - dim + d)
    return x
