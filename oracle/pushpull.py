"""ORACLE — test infrastructure only.

numpy front-end of the C restatement in `pushpull_oracle.c`, mirroring the
reference's `nitorch/_C/src/pushpull.cpp` entry points (argument meaning,
output shapes, which outputs the backward passes return).

Arrays follow the reference's layouts: input `[N, C, *spatial]`, grid
`[N, *spatial_out, dim]`.  `bound` / `order` are lists padded to three
entries by repeating the last one (pushpull_common.cpp:96-110); bound codes
are the reference's BoundType values (bounds.h:7-17).

`oob_f32=True` makes the FP64 restatement decide which samples are out of
bounds the way the reference does for float32 data: in float arithmetic,
`x >= -tol && x < (float)(n - 1) + tol` with `tol = (float)TINY`
(bounds_common.h:168-171, pushpull_common.cpp:900).  For float32
coordinates that sit exactly on an edge (e.g. float32(44.05) with n = 45)
the FP64 test and the float test disagree; weights and sums stay FP64.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

BOUND_CODES = dict(replicate=0, dct1=1, dct2=2, dst1=3, dst2=4, dft=5,
                   sliding=6, zero=7, nocheck=8)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for sfx, real in (("_f64", ctypes.c_double), ("_f32", ctypes.c_float)):
            fn = getattr(_lib, "or_pushpull" + sfx)
            fn.restype = ctypes.c_int
            for name in ("or_spline_weight", "or_spline_grad", "or_spline_hess"):
                f = getattr(_lib, name + sfx)
                f.restype = ctypes.c_double
                f.argtypes = [ctypes.c_int, ctypes.c_double]
            f = getattr(_lib, "or_bound_index" + sfx)
            f.restype = ctypes.c_int64
            f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64]
            f = getattr(_lib, "or_bound_sign" + sfx)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64]
    return _lib


def pad3(x):
    x = list(x) if isinstance(x, (list, tuple)) else [x]
    return (x + [x[-1]] * 3)[:3]


def bound_codes(bound):
    return [BOUND_CODES[b] if isinstance(b, str) else int(b) for b in pad3(bound)]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _call(dtype, dim, bound, order, extrapolate, flags, N, C, src_size,
          trgt_size, trgt_K, src, trgt, grid, out, gout, quirk_zorder, oob_f32=False):
    sfx = "_f64" if dtype == np.float64 else "_f32"
    fn = getattr(lib(), "or_pushpull" + sfx)
    b = (ctypes.c_int * 3)(*bound_codes(bound))
    o = (ctypes.c_int * 3)(*[int(v) for v in pad3(order)])
    ss = (ctypes.c_int64 * 3)(*(list(src_size) + [1, 1])[:3])
    ts = (ctypes.c_int64 * 3)(*(list(trgt_size) + [1, 1])[:3])
    rc = fn(ctypes.c_int(dim), b, o, ctypes.c_int(int(extrapolate)),
            *[ctypes.c_int(int(f)) for f in flags],
            ctypes.c_int64(N), ctypes.c_int64(C), ss, ts, ctypes.c_int(trgt_K),
            _ptr(src), _ptr(trgt), _ptr(grid), _ptr(out), _ptr(gout),
            ctypes.c_int(int(bool(quirk_zorder)) | (2 if oob_f32 else 0)))
    if rc != 0:
        raise RuntimeError("oracle pushpull failed")


def _prep(*arrs, dtype):
    return [None if a is None else np.ascontiguousarray(a, dtype=dtype) for a in arrs]


def grid_pull(input, grid, bound, order, extrapolate=0, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:75-107 — returns [N, C, *grid_spatial]."""
    dtype = input.dtype
    input, grid = _prep(input, grid, dtype=dtype)
    dim = grid.shape[-1]
    N, C = input.shape[:2]
    out = np.empty((N, C) + grid.shape[1:-1], dtype)
    _call(dtype, dim, bound, order, extrapolate, (1, 0, 0, 0, 0), N, C,
          input.shape[2:], grid.shape[1:-1], 0, input, None, grid, out, None,
          quirk_zorder, oob_f32)
    return out


def grid_grad(input, grid, bound, order, extrapolate=0, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:268-300 — returns [N, C, *grid_spatial, dim]."""
    dtype = input.dtype
    input, grid = _prep(input, grid, dtype=dtype)
    dim = grid.shape[-1]
    N, C = input.shape[:2]
    out = np.empty((N, C) + grid.shape[1:-1] + (dim,), dtype)
    _call(dtype, dim, bound, order, extrapolate, (0, 0, 0, 0, 1), N, C,
          input.shape[2:], grid.shape[1:-1], 0, input, None, grid, out, None,
          quirk_zorder, oob_f32)
    return out


def grid_push(input, grid, shape, bound, order, extrapolate=0, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:131-184 — input [N, C, *grid_spatial] -> [N, C, *shape]."""
    dtype = input.dtype
    input, grid = _prep(input, grid, dtype=dtype)
    dim = grid.shape[-1]
    N, C = input.shape[:2]
    shape = tuple(shape) if shape is not None else tuple(input.shape[2:])
    out = np.empty((N, C) + shape, dtype)
    _call(dtype, dim, bound, order, extrapolate, (0, 1, 0, 0, 0), N, C,
          shape, grid.shape[1:-1], 0, None, input, grid, out, None,
          quirk_zorder, oob_f32)
    return out


def grid_count(grid, shape, bound, order, extrapolate=0, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:204-247 — returns [N, 1, *shape]."""
    dtype = grid.dtype
    (grid,) = _prep(grid, dtype=dtype)
    dim = grid.shape[-1]
    N = grid.shape[0]
    shape = tuple(shape) if shape is not None else tuple(grid.shape[1:-1])
    out = np.empty((N, 1) + shape, dtype)
    _call(dtype, dim, bound, order, extrapolate, (0, 0, 1, 0, 0), N, 1,
          shape, grid.shape[1:-1], 0, None, None, grid, out, None,
          quirk_zorder, oob_f32)
    return out


def grid_pull_backward(grad, input, grid, bound, order, extrapolate=0,
                       need_input=True, need_grid=True, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:109-128 — (grad_input, grad_grid); None where not needed."""
    dtype = input.dtype
    grad, input, grid = _prep(grad, input, grid, dtype=dtype)
    dim = grid.shape[-1]
    N, C = input.shape[:2]
    gi = np.empty(input.shape, dtype) if need_input else None
    gg = np.empty(grid.shape, dtype) if need_grid else None
    _call(dtype, dim, bound, order, extrapolate,
          (0, int(need_input), 0, int(need_grid), 0), N, C,
          input.shape[2:], grid.shape[1:-1], 0, input, grad, grid, gi, gg,
          quirk_zorder, oob_f32)
    return gi, gg


def grid_push_backward(grad, input, grid, bound, order, extrapolate=0,
                       need_input=True, need_grid=True, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:186-202 — source=grad (push output space), target=input."""
    dtype = input.dtype
    grad, input, grid = _prep(grad, input, grid, dtype=dtype)
    dim = grid.shape[-1]
    N, C = input.shape[:2]
    gi = np.empty(input.shape, dtype) if need_input else None
    gg = np.empty(grid.shape, dtype) if need_grid else None
    _call(dtype, dim, bound, order, extrapolate,
          (int(need_input), 0, 0, int(need_grid), 0), N, C,
          grad.shape[2:], grid.shape[1:-1], 0, grad, input, grid, gi, gg,
          quirk_zorder, oob_f32)
    return gi, gg


def grid_count_backward(grad, grid, bound, order, extrapolate=0, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:249-265 — grid gradient of count; source=grad, no target."""
    dtype = grid.dtype
    grad, grid = _prep(grad, grid, dtype=dtype)
    dim = grid.shape[-1]
    N = grid.shape[0]
    gg = np.empty(grid.shape, dtype)
    _call(dtype, dim, bound, order, extrapolate, (0, 0, 0, 1, 0), N,
          grad.shape[1], grad.shape[2:], grid.shape[1:-1], 0, grad, None,
          grid, None, gg, quirk_zorder, oob_f32)
    return gg


def grid_grad_backward(grad, input, grid, bound, order, extrapolate=0,
                       need_input=True, need_grid=True, quirk_zorder=False, oob_f32=False):
    """pushpull.cpp:303-327 — target = grad [N, C, *out, dim] (trgt_K = dim)."""
    dtype = input.dtype
    grad, input, grid = _prep(grad, input, grid, dtype=dtype)
    dim = grid.shape[-1]
    N, C = input.shape[:2]
    gi = np.empty(input.shape, dtype) if need_input else None
    gg = np.empty(grid.shape, dtype) if need_grid else None
    _call(dtype, dim, bound, order, extrapolate,
          (0, int(need_input), 0, int(need_grid), 0), N, C,
          input.shape[2:], grid.shape[1:-1], dim, input, grad, grid, gi, gg,
          quirk_zorder, oob_f32)
    return gi, gg


def spline_weight(order, x, f32=False):
    return getattr(lib(), "or_spline_weight" + ("_f32" if f32 else "_f64"))(order, x)


def spline_grad(order, x):
    return lib().or_spline_grad_f64(order, x)


def spline_hess(order, x):
    return lib().or_spline_hess_f64(order, x)


def bound_index(b, i, n):
    return lib().or_bound_index_f64(BOUND_CODES.get(b, b) if isinstance(b, str) else b, i, n)


def bound_sign(b, i, n):
    return lib().or_bound_sign_f64(BOUND_CODES.get(b, b) if isinstance(b, str) else b, i, n)
