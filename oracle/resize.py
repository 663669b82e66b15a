"""ORACLE — test infrastructure only.

FP64 numpy restatement of the reference's compiled resize
(nitorch/_C/src/resize.cpp:25-59; nitorch/_C/src/impl/resize_common.cpp):

* output allocation (prepare_tensors :1490-1545): pull -> floor(n * f),
  adjoint -> source ceil(n / f) (the factor is inverted for the adjoint);
* per-axis shift / scale (prepare_affine :1548-1567, GridAlignType
  edge 0, center 1, first 2, last 3);
* target voxel w samples the source at x = scale * w + shift with the
  order's B-spline and the pull's boundary folding, never masked as out of
  bounds (a pull with extrapolation, here the oracle's own pull / push of
  `pushpull_oracle.c`); the adjoint splats into a zeroed source;
* a given output is accumulated into, except by the isotropic nearest and
  linear kernels (any dim) and the isotropic quadratic 3-D kernel, which
  overwrite it (resize3d_nearest :1383, resize3d_linear :1208,
  resize3d_quadratic; dispatch :440-530); isotropic nearest rounds
  scale * w + shift in double (:1356-1358);
* `normalize` multiplies by the scales with the fall-through switch of
  :1640-1646 (1-D: s0 * s0 s1 * s0 s1 s2, 2-D: s0 s1 * s0 s1 s2, 3-D: s0 s1 s2).
Arrays are channel-first: [N, C, *spatial], 1 to 3 spatial dims.
"""
import math

import numpy as np

from . import pushpull as PP

EDGE, CENTER, FIRST, LAST = 0, 1, 2, 3
_MODES = {"e": EDGE, "c": CENTER, "f": FIRST, "l": LAST}


def mode_code(m):
    return _MODES[m[0].lower()] if isinstance(m, str) else int(m)


def _unfold(v, default):
    v = list(v) if isinstance(v, (list, tuple)) else ([v] if v is not None else [])
    if not v:
        v = [default]
    return (v + [v[-1]] * 3)[:3]


def affine(f, Ds, Dt, mode):
    """(shift, scale) of one axis (resize_common.cpp:1548-1567)."""
    shift, scale = 0., 1. / f
    if mode == EDGE:
        shift, scale = 0.5 * (Ds / Dt - 1), Ds / Dt
    elif mode == CENTER:
        shift = 0.
        scale = (Ds - 1) / (Dt - 1) if Dt != 1 else (math.nan if Ds == 1 else math.copysign(math.inf, Ds - 1))
    elif mode == LAST:
        shift = f * (Ds - 1) / (Dt - 1) if Dt != 1 else (math.nan if Ds == 1 else math.copysign(math.inf, f * (Ds - 1)))
    return shift, scale


def out_shape(in_shape, factor, adjoint):
    f = _unfold(factor, 1.)
    if adjoint:
        f = [1. / v for v in f]
        return [int(math.ceil(n * v)) for n, v in zip(in_shape, f)]
    return [int(math.floor(n * v)) for n, v in zip(in_shape, f)]


def resize(x, factor=None, bound="dct2", order=1, mode=CENTER, shape=None, out=None,
           adjoint=False, normalize=False):
    """ni::resize(input, output, factor, bound, interpolation, mode,
    do_adjoint, normalize) on float64 arrays; returns the output."""
    x = np.asarray(x, np.float64)
    dim = x.ndim - 2
    f = _unfold(factor, 1.)
    bound = _unfold(bound, "dct2")
    order = [int(o) for o in _unfold(order, 1)]
    mode = [mode_code(m) for m in _unfold(mode, CENTER)]
    given = out is not None
    if out is None:
        shp = list(shape) if shape is not None else out_shape(x.shape[2:], f, adjoint)
        out = np.zeros(x.shape[:2] + tuple(shp))
    else:
        out = np.array(out, np.float64, copy=True)
    src, tgt = (out, x) if adjoint else (x, out)
    fa = [1. / v for v in f] if adjoint else f
    shifts, scales = [], []
    for d in range(3):
        Ds = src.shape[2 + d] if d < dim else 1
        Dt = tgt.shape[2 + d] if d < dim else 1
        s, c = affine(fa[d], float(Ds), float(Dt), mode[d])
        shifts.append(s)
        scales.append(c)
    iso = order[0] == order[1] == order[2]
    # sampling coordinates of the target lattice in source voxels
    axes = [np.arange(tgt.shape[2 + d], dtype=np.float64) * scales[d] + shifts[d] for d in range(dim)]
    if iso and order[0] == 0:
        # std::round of the double coordinate: nearest, exact ties away from zero
        axes = [np.where(np.abs(a - np.trunc(a)) == 0.5, np.trunc(a) + np.sign(a), np.round(a))
                for a in axes]
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1)
    grid = np.ascontiguousarray(np.broadcast_to(grid[None], (x.shape[0],) + grid.shape))
    if not adjoint:
        y = PP.grid_pull(src, grid, bound[:dim], order[:dim], 1)
        overwrite = iso and (order[0] <= 1 or (order[0] == 2 and dim == 3))
        res = y if (overwrite or not given) else tgt + y
    else:
        res = PP.grid_push(tgt, grid, src.shape[2:], bound[:dim], order[:dim], 1)
    if normalize:
        m = scales[0] * scales[1] * scales[2]
        if dim == 2:
            m *= scales[0] * scales[1]
        if dim == 1:
            m *= scales[0] * (scales[0] * scales[1])
        res = res * m
    return res
