"""ORACLE — test infrastructure only.

numpy front-end of `solver_oracle.c` (FP64) for 3-D displacement fields in
the reference's C++ layout [N, 3, X, Y, Z]; Hessians [N, CC, X, Y, Z] with
CC in {1, 3, 6} (6 = h00 h11 h22 h01 h02 h12).  `weight` (optional, every
entry point): voxel-wise RLS weights [N, 1, X, Y, Z] (or [N, X, Y, Z]) of the
absolute / membrane energies.
"""
import ctypes

import numpy as np

from .pushpull import lib as _lib, bound_codes

_c = ctypes


def _setup():
    L = _lib()
    if getattr(L, "_solver_ready", False):
        return L
    D, I64, I = _c.POINTER(_c.c_double), _c.c_int64, _c.c_int
    PI, PI64 = _c.POINTER(_c.c_int), _c.POINTER(_c.c_int64)
    L.or_regulariser_grid.argtypes = [D, D, D, I, I64, I64, I64, I64, PI, D] + [_c.c_double] * 5 + [D]
    L.or_regulariser_grid.restype = None
    L.or_relax_grid.argtypes = [D, D, I, D, I64, I64, I64, I64, PI, D] + [_c.c_double] * 5 + [I, I, D]
    L.or_relax_grid.restype = None
    L.or_prolong.argtypes = [D, D, I64, I64, PI64, PI64, PI]
    L.or_prolong.restype = None
    L.or_restrict.argtypes = [D, D, I64, I64, PI64, PI64, PI]
    L.or_restrict.restype = None
    L.or_fmg_grid.argtypes = [D, D, D, I, I64, I64, I64, I64, PI, D] + [_c.c_double] * 5 + [I, I, I, I, I, I, D]
    L.or_fmg_grid.restype = None
    L.or_cg_grid.argtypes = [D, D, D, I, I64, I64, I64, I64, PI, D] + [_c.c_double] * 5 + [I, _c.c_double, D]
    L.or_cg_grid.restype = I
    L.or_pcg_grid.argtypes = [D, D, D, I, I64, I64, I64, I64, PI, D] + [_c.c_double] * 5 + [I, _c.c_double, D]
    L.or_pcg_grid.restype = I
    L.or_precond_grid.argtypes = [D, D, D, I, I64, I64, I64, I64, PI, D] + [_c.c_double] * 5 + [D]
    L.or_precond_grid.restype = None
    L.or_relax_plan.argtypes = [I64, I64, I64, PI, I, PI]
    L.or_relax_plan.restype = I
    L._solver_ready = True
    return L


def _d(a):
    return a.ctypes.data_as(_c.POINTER(_c.c_double)) if a is not None else None


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _w(weight, shape):
    """RLS weights as a contiguous [N, 1, X, Y, Z] float64 array (or None)."""
    if weight is None:
        return None
    w = np.asarray(weight, dtype=np.float64)
    if w.ndim == 4:
        w = w[:, None]
    return np.ascontiguousarray(np.broadcast_to(w, (shape[0], 1) + tuple(shape[2:])))


def _vs(voxel_size):
    v = list(np.atleast_1d(voxel_size).astype(float))
    v = (v + [v[-1]] * 3)[:3]
    return (_c.c_double * 3)(*v)


def _b(bound):
    return (_c.c_int * 3)(*bound_codes(bound))


def _pen(absolute, membrane, bending, lame):
    lame = list(np.atleast_1d(lame).astype(float)) if lame is not None else [0.0]
    lame = (lame + [lame[-1]])[:2]
    return [float(absolute), float(membrane), float(bending), lame[0], lame[1]]


def regulariser_grid(x, hessian=None, absolute=0, membrane=0, bending=0, lame=0,
                     voxel_size=1, bound="dft", weight=None):
    """(H + L) x for x [N, 3, X, Y, Z]."""
    L = _setup()
    x, h = _f64(x), _f64(hessian)
    w = _w(weight, x.shape)
    out = np.empty_like(x)
    N, _, X, Y, Z = x.shape
    CC = 0 if h is None else h.shape[1]
    L.or_regulariser_grid(_d(x), _d(out), _d(h), CC, N, X, Y, Z, _b(bound), _vs(voxel_size),
                          *_pen(absolute, membrane, bending, lame), _d(w))
    return out


def relax_grid(g, hessian=None, x=None, absolute=0, membrane=0, bending=0, lame=0,
               voxel_size=1, bound="dft", nb_iter=1, quirk_fold=False, product_order=False, weight=None,
               reverse_within_classes=False):
    """Gauss-Seidel sweeps.  quirk_fold: the reference's fold order and sign
    typos; product_order: the colour classes the HIP kernels use (see
    relax_plan); reverse_within_classes (with product_order): visit each
    class's voxels in reverse order, which changes nothing exactly when
    every class is an independent set of the stencil."""
    L = _setup()
    g, h = _f64(g), _f64(hessian)
    x = np.zeros_like(g) if x is None else _f64(x).copy()
    N, _, X, Y, Z = g.shape
    CC = 0 if h is None else h.shape[1]
    mode = (3 if reverse_within_classes else 2) if product_order else int(bool(quirk_fold))
    w = _w(weight, g.shape)
    L.or_relax_grid(_d(g), _d(h), CC, _d(x), N, X, Y, Z, _b(bound), _vs(voxel_size),
                    *_pen(absolute, membrane, bending, lame), int(nb_iter), mode, _d(w))
    return x


def relax_plan(shape, bound="dft", band=3):
    """The product's colour classes for a lattice: ("linear", m, (a, b, c))
    -- colour (a x + b y + c z) mod m, classes 0 .. m-1 in turn -- or
    ("axes", n_classes, None), the per-axis period classes.  band: 3 with a
    bending energy, 2 with Lame only, 0 membrane only, 1 absolute only."""
    L = _setup()
    abc = (_c.c_int * 3)()
    m = L.or_relax_plan(*[int(s) for s in shape], _b(bound), int(band), abc)
    if m:
        return "linear", m, tuple(abc)
    return "axes", abc[0], None


def prolong(x, out_shape, bound="dct2"):
    L = _setup()
    x = _f64(x)
    N, C = x.shape[:2]
    out = np.empty((N, C) + tuple(out_shape))
    L.or_prolong(_d(x), _d(out), N, C, (_c.c_int64 * 3)(*x.shape[2:]), (_c.c_int64 * 3)(*out_shape), _b(bound))
    return out


def restrict(x, out_shape=None, bound="dct2"):
    L = _setup()
    x = _f64(x)
    N, C = x.shape[:2]
    if out_shape is None:
        out_shape = tuple(1 + (s - 1) // 2 for s in x.shape[2:])
    out = np.empty((N, C) + tuple(out_shape))
    L.or_restrict(_d(x), _d(out), N, C, (_c.c_int64 * 3)(*x.shape[2:]), (_c.c_int64 * 3)(*out_shape), _b(bound))
    return out


def fmg_grid(hessian, g, x=None, absolute=0, membrane=0, bending=0, lame=0, voxel_size=1,
             bound="dft", nb_cycles=2, nb_iter=2, max_levels=16, quirk_fold=False, quirk_zscale=False,
             product_order=False, use_cg=False, weight=None):
    L = _setup()
    g, h = _f64(g), _f64(hessian)
    x = np.zeros_like(g) if x is None else _f64(x).copy()
    N, _, X, Y, Z = g.shape
    CC = 0 if h is None else h.shape[1]
    w = _w(weight, g.shape)
    L.or_fmg_grid(_d(h), _d(g), _d(x), CC, N, X, Y, Z, _b(bound), _vs(voxel_size),
                  *_pen(absolute, membrane, bending, lame), int(nb_cycles), int(nb_iter),
                  int(max_levels), 2 if product_order else int(bool(quirk_fold)), int(quirk_zscale),
                  int(bool(use_cg)), _d(w))
    return x


def precond_grid(g, hessian=None, absolute=0, membrane=0, bending=0, lame=0, voxel_size=1,
                 bound="dft", weight=None):
    """(H (1+1e-6) + diag(L) (1+1e-6))^-1 g, voxel by voxel (with RLS
    weights the diagonal of the weighted operator, which reads the
    neighbours' weights through `bound`)."""
    L = _setup()
    g, h = _f64(g), _f64(hessian)
    out = np.empty_like(g)
    N, _, X, Y, Z = g.shape
    CC = 0 if h is None else h.shape[1]
    w = _w(weight, g.shape)
    L.or_precond_grid(_d(h), _d(g), _d(out), CC, N, X, Y, Z, _b(bound), _vs(voxel_size),
                      *_pen(absolute, membrane, bending, lame), _d(w))
    return out


def pcg_grid(hessian, g, x=None, absolute=0, membrane=0, bending=0, lame=0, voxel_size=1,
             bound="dft", nb_iter=16, tol=0.0, weight=None):
    """Preconditioned CG as the reference's do_pcg; returns (x, iterations)."""
    L = _setup()
    g, h = _f64(g), _f64(hessian)
    x = np.zeros_like(g) if x is None else _f64(x).copy()
    N, _, X, Y, Z = g.shape
    CC = 0 if h is None else h.shape[1]
    w = _w(weight, g.shape)
    it = L.or_pcg_grid(_d(h), _d(g), _d(x), CC, N, X, Y, Z, _b(bound), _vs(voxel_size),
                       *_pen(absolute, membrane, bending, lame), int(nb_iter), float(tol), _d(w))
    return x, it


def cg_grid(hessian, g, x=None, absolute=0, membrane=0, bending=0, lame=0, voxel_size=1,
            bound="dft", max_iter=10000, tol=1e-12, weight=None):
    L = _setup()
    g, h = _f64(g), _f64(hessian)
    x = np.zeros_like(g) if x is None else _f64(x).copy()
    N, _, X, Y, Z = g.shape
    CC = 0 if h is None else h.shape[1]
    w = _w(weight, g.shape)
    it = L.or_cg_grid(_d(h), _d(g), _d(x), CC, N, X, Y, Z, _b(bound), _vs(voxel_size),
                      *_pen(absolute, membrane, bending, lame), int(max_iter), float(tol), _d(w))
    return x, it
