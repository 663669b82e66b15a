"""Displacement-field solver front-end, drop-in for the grid functions of
`nitorch/_C/solve.py` (c_regulariser_grid :100-140, c_relax_grid :203-270,
c_precond_grid :324-384, c_pcg_grid :445-505, c_fmg_grid :574-641) and
`nitorch.spatial.solve_grid_fmg`
(nitorch/spatial/_solvers.py:978-1039), running the HIP kernels of
libnitorch_amd.so.

Shapes follow the reference's Python layer: fields `(N, *spatial, 3)`,
Hessians `(N, *spatial, CC)` with CC in {1, 3, 6} (channel last); 3-D only.
Penalties: absolute, membrane, bending, lame=(shear, div), all multiplied by
`factor`.  Voxel-wise weights (RLS) `weight` of shape (N, *spatial) (or
(N, *spatial, 1)) weight the absolute and membrane energies, as the
reference's `weight` argument (bending / Lame penalties with weights raise,
as there).
"""
import ctypes

import torch

from . import _lib as L
from .grid import bound_to_nitorch, make_list

L.register_signatures({
    "ni_regulariser_grid": [L._I32, L._I64, L._PI64, L._P, L._P, L._P, L._I32, L._P, L._PI32,
                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), L._P],
    "ni_residual_grid": [L._I32, L._I64, L._PI64, L._P, L._P, L._P, L._P, L._I32, L._P, L._PI32,
                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), L._P, L._P],
    "ni_relax_grid": [L._I32, L._I64, L._PI64, L._P, L._I32, L._P, L._P, L._P, L._PI32,
                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), L._I32, L._P],
    "ni_fmg_restriction": [L._I32, L._I64, L._I64, L._PI64, L._PI64, L._P, L._P, L._PI32, L._P],
    "ni_fmg_prolongation": [L._I32, L._I64, L._I64, L._PI64, L._PI64, L._P, L._P, L._PI32, L._P],
    "ni_fmg_grid": [L._I32, L._I64, L._PI64, L._P, L._I32, L._P, L._P, L._P, L._PI32,
                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                    L._I32, L._I32, L._I32, L._I32, L._I32, L._P, L._I64, L._P],
    "ni_fmg_solve": [L._I32, L._I64, L._PI64, L._P, L._I32, L._P, L._P, L._P, L._PI32,
                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                     L._I32, L._I32, L._I32, L._I32, ctypes.c_double, L._I32, L._P, L._I64,
                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32), L._P],
    "ni_pcg_grid": [L._I32, L._I64, L._PI64, L._P, L._I32, L._P, L._P, L._P, L._PI32,
                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                    L._I32, ctypes.c_double, L._P, L._I64, ctypes.POINTER(ctypes.c_int32), L._P],
    "ni_precond_grid": [L._I32, L._I64, L._PI64, L._P, L._I32, L._P, L._P, L._P, L._PI32,
                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), L._P],
})


def _ws_fn():
    fn = L.lib().ni_fmg_grid_workspace
    fn.restype = ctypes.c_int64
    fn.argtypes = [L._I32, L._I64, L._PI64, L._I32, L._I32, L._I32, L._I32]
    return fn


def _pcg_ws_fn():
    fn = L.lib().ni_pcg_grid_workspace
    fn.restype = ctypes.c_int64
    fn.argtypes = [L._I32, L._I64, L._PI64]
    return fn


def _dvec(vals):
    vals = [float(v) for v in vals]
    return (ctypes.c_double * len(vals))(*vals)


def _penalty(absolute, membrane, bending, lame, factor):
    lame = make_list(lame) if lame is not None else [0.0]
    lame = (lame + [lame[-1]])[:2]
    return [absolute * factor, membrane * factor, bending * factor,
            lame[0] * factor, lame[1] * factor]


def _voxel(voxel_size):
    vs = make_list(voxel_size) if voxel_size is not None else [1.0]
    vs = [float(v) for v in vs] or [1.0]
    return _dvec((vs + [vs[-1]] * 3)[:3])


def _bounds(bound):
    b = [int(x) for x in bound_to_nitorch(make_list(bound), "int")]
    return L.i32arr((b + [b[-1]] * 3)[:3])


def _to_cf(t, name, dim=3):
    """(N, *spatial, C) -> contiguous (N, C, *spatial)."""
    if t is None:
        return None
    while t.dim() < dim + 2:
        t = t.unsqueeze(0)
    return t.movedim(-1, 1).contiguous()


def _check(gradient, hessian):
    if gradient.shape[-1] != 3 or gradient.dim() < 4:
        raise ValueError("nitorch_amd solvers handle 3-D displacement fields (..., X, Y, Z, 3)")
    L.require_device(gradient, hessian)
    L.dtype_code(gradient.dtype)
    if hessian is not None:
        if hessian.shape[-1] not in (1, 3, 6):
            raise ValueError(f"hessian must have 1, 3 or 6 channels, got {hessian.shape[-1]}")
        if hessian.dtype != gradient.dtype or hessian.device != gradient.device:
            raise ValueError("hessian and gradient must share dtype and device")


def _weight_cf(weight, like):
    """RLS weights of the reference's grid layer -- (N, *spatial), or
    (N, *spatial, 1) -- as a contiguous (N, 1, *spatial) tensor broadcast to
    the batch of `like` (channel-first field)."""
    if weight is None:
        return None
    w = torch.as_tensor(weight, dtype=like.dtype, device=like.device)
    dim = like.dim() - 2
    if w.dim() == dim + 2 and w.shape[-1] == 1:
        w = w[..., 0]
    while w.dim() < dim + 1:
        w = w.unsqueeze(0)
    w = w.unsqueeze(1)
    return w.expand((like.shape[0], 1) + tuple(like.shape[2:])).contiguous()


# ---------------------------------------------------------------------------
# channel-first primitives (the reference's C++ layer)
# ---------------------------------------------------------------------------
def regulariser_grid_cf(x, hessian=None, penalty=(0, 0, 0, 0, 0), voxel_size=(1, 1, 1),
                        bound="dft", out=None, weight=None):
    """out = (H + L) x for contiguous x [N, 3, X, Y, Z], H [N, CC, X, Y, Z],
    RLS weights [N, 1, X, Y, Z]."""
    out = torch.empty_like(x) if out is None else out
    CC = 0 if hessian is None else hessian.shape[1]
    L.check(L.lib().ni_regulariser_grid(
        L.dtype_code(x.dtype), x.shape[0], L.i64arr(x.shape[2:]), L.ptr(x), L.ptr(out),
        L.ptr(hessian), CC, L.ptr(weight), _bounds(bound), _voxel(voxel_size), _dvec(penalty),
        L.stream_of(x)), "regulariser_grid")
    return out


def residual_grid_cf(x, g, hessian=None, penalty=(0, 0, 0, 0, 0), voxel_size=(1, 1, 1),
                     bound="dft", out=None, norms=None, weight=None):
    """r = g - (H + L) x; if `norms` (float64 device tensor of 2) is given it
    receives ||r||^2 and ||g||^2."""
    out = torch.empty_like(x) if out is None else out
    CC = 0 if hessian is None else hessian.shape[1]
    L.check(L.lib().ni_residual_grid(
        L.dtype_code(x.dtype), x.shape[0], L.i64arr(x.shape[2:]), L.ptr(x), L.ptr(g), L.ptr(out),
        L.ptr(hessian), CC, L.ptr(weight), _bounds(bound), _voxel(voxel_size), _dvec(penalty),
        L.ptr(norms), L.stream_of(x)), "residual_grid")
    return out


def relax_colouring(shape, bound="dft", penalty=(0, 0, 0, 0, 0)):
    """Colour classes of the Gauss-Seidel sweeps on a 3-D lattice (host
    only): ("linear", m, (a, b, c)) -- colour (a x + b y + c z) mod m,
    classes 0 .. m-1 in turn -- or ("axes", n_classes, None) for per-axis
    period classes."""
    fn = L.lib().ni_relax_colouring
    fn.restype = ctypes.c_int
    fn.argtypes = [L._PI64, L._PI32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]
    plan = (ctypes.c_int32 * 5)()
    L.check(fn(L.i64arr(shape), _bounds(bound), _dvec(penalty), plan), "relax_colouring")
    if plan[0]:
        return "linear", plan[1], (plan[2], plan[3], plan[4])
    return "axes", plan[1], None


def relax_grid_cf(g, x, hessian=None, penalty=(0, 0, 0, 0, 0), voxel_size=(1, 1, 1),
                  bound="dft", nb_iter=2, weight=None):
    """nb_iter Gauss-Seidel sweeps, x updated in place."""
    CC = 0 if hessian is None else hessian.shape[1]
    L.check(L.lib().ni_relax_grid(
        L.dtype_code(g.dtype), g.shape[0], L.i64arr(g.shape[2:]), L.ptr(hessian), CC, L.ptr(weight),
        L.ptr(g), L.ptr(x), _bounds(bound), _voxel(voxel_size), _dvec(penalty), int(nb_iter),
        L.stream_of(g)), "relax_grid")
    return x


def fmg_restriction(input, output=None, bound="dct2"):
    """nitorch/_C/src/multires.h fmg_restriction: [N, C, *s] -> [N, C, *ceil(s/2)]."""
    L.require_device(input)
    input = input.contiguous()
    if output is None:
        output = input.new_empty(tuple(input.shape[:2]) + tuple(1 + (s - 1) // 2 for s in input.shape[2:]))
    L.check(L.lib().ni_fmg_restriction(
        L.dtype_code(input.dtype), input.shape[0], input.shape[1], L.i64arr(input.shape[2:]),
        L.i64arr(output.shape[2:]), L.ptr(input), L.ptr(output), _bounds(bound), L.stream_of(input)),
        "fmg_restriction")
    return output


def fmg_prolongation(input, output=None, bound="dct2"):
    """nitorch/_C/src/multires.h fmg_prolongation: [N, C, *s] -> [N, C, *2s]."""
    L.require_device(input)
    input = input.contiguous()
    if output is None:
        output = input.new_empty(tuple(input.shape[:2]) + tuple(2 * s for s in input.shape[2:]))
    L.check(L.lib().ni_fmg_prolongation(
        L.dtype_code(input.dtype), input.shape[0], input.shape[1], L.i64arr(input.shape[2:]),
        L.i64arr(output.shape[2:]), L.ptr(input), L.ptr(output), _bounds(bound), L.stream_of(input)),
        "fmg_prolongation")
    return output


def fmg_grid_cf(hessian, g, x=None, penalty=(0, 0, 0, 0, 0), voxel_size=(1, 1, 1), bound="dft",
                nb_cycles=2, nb_iter=2, max_levels=16, extra_vcycles=0, workspace=None,
                use_cg=False, weight=None):
    """Full multigrid on contiguous channel-first tensors; x updated in place.
    use_cg: PCG smoothing (the reference's solver='cg'), else Gauss-Seidel."""
    x = torch.zeros_like(g) if x is None else x
    CC = 0 if hessian is None else hessian.shape[1]
    dt = L.dtype_code(g.dtype)
    nbytes = _ws_fn()(dt, g.shape[0], L.i64arr(g.shape[2:]), CC, int(max_levels), int(bool(use_cg)),
                      int(weight is not None))
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=g.device)
    L.check(L.lib().ni_fmg_grid(
        dt, g.shape[0], L.i64arr(g.shape[2:]), L.ptr(hessian), CC, L.ptr(weight), L.ptr(g), L.ptr(x),
        _bounds(bound), _voxel(voxel_size), _dvec(penalty), int(nb_cycles), int(nb_iter),
        int(max_levels), int(bool(use_cg)), int(extra_vcycles), L.ptr(workspace), int(nbytes),
        L.stream_of(g)), "fmg_grid")
    return x


def pcg_grid_cf(hessian, g, x=None, penalty=(0, 0, 0, 0, 0), voxel_size=(1, 1, 1), bound="dft",
                nb_iter=16, tol=0.0, workspace=None, weight=None):
    """Block-Jacobi preconditioned CG (nitorch/_C/src/impl/pcg_common.cpp:38-85) on
    contiguous channel-first tensors; x updated in place.  Returns (x, iterations)."""
    x = torch.zeros_like(g) if x is None else x
    CC = 0 if hessian is None else hessian.shape[1]
    dt = L.dtype_code(g.dtype)
    nbytes = _pcg_ws_fn()(dt, g.shape[0], L.i64arr(g.shape[2:]))
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=g.device)
    iters = ctypes.c_int32(0)
    L.check(L.lib().ni_pcg_grid(
        dt, g.shape[0], L.i64arr(g.shape[2:]), L.ptr(hessian), CC, L.ptr(weight), L.ptr(g), L.ptr(x),
        _bounds(bound), _voxel(voxel_size), _dvec(penalty), int(nb_iter), float(tol),
        L.ptr(workspace), int(nbytes), ctypes.byref(iters), L.stream_of(g)), "pcg_grid")
    return x, iters.value


def precond_grid_cf(hessian, g, penalty=(0, 0, 0, 0, 0), voxel_size=(1, 1, 1), bound="dft", out=None,
                    weight=None):
    """(H + diag L)^-1 g voxel by voxel (precond_grid_common.cpp:620-660)."""
    out = torch.empty_like(g) if out is None else out
    CC = 0 if hessian is None else hessian.shape[1]
    L.check(L.lib().ni_precond_grid(
        L.dtype_code(g.dtype), g.shape[0], L.i64arr(g.shape[2:]), L.ptr(hessian), CC, L.ptr(weight), L.ptr(g),
        L.ptr(out), _bounds(bound), _voxel(voxel_size), _dvec(penalty), L.stream_of(g)),
        "precond_grid")
    return out


# ---------------------------------------------------------------------------
# reference-shaped entry points (channel last)
# ---------------------------------------------------------------------------
def regulariser_grid(input, weight=None, hessian=None, absolute=0, membrane=0, bending=0,
                     lame=0, factor=1, voxel_size=1, bound="dft", output=None):
    """(hessian + regulariser) @ input for input (N, *shape, 3)
    (nitorch/_C/solve.py c_regulariser_grid)."""
    _check(input, hessian)
    pad = 5 - input.dim()
    x = _to_cf(input, "input")
    h = _to_cf(hessian, "hessian")
    out = regulariser_grid_cf(x, h, _penalty(absolute, membrane, bending, lame, factor),
                              voxel_size, bound, weight=_weight_cf(weight, x))
    out = out.movedim(1, -1)
    if output is not None:
        output.copy_(out.reshape(output.shape))
        out = output
    for _ in range(max(pad, 0)):
        out = out.squeeze(0)
    return out


def relax_grid(gradient, weight=None, hessian=None, absolute=0, membrane=0, bending=0, lame=0,
               factor=1, voxel_size=1, bound="dft", nb_iter=2, output=None):
    """Gauss-Seidel relaxation of (H + L) x = g (nitorch/_C/solve.py c_relax_grid);
    `output` is the initial guess and receives the result."""
    _check(gradient, hessian)
    pad = 5 - gradient.dim()
    g = _to_cf(gradient, "gradient")
    h = _to_cf(hessian, "hessian")
    x = torch.zeros_like(g) if output is None else _to_cf(output, "output")
    relax_grid_cf(g, x, h, _penalty(absolute, membrane, bending, lame, factor), voxel_size,
                  bound, nb_iter, weight=_weight_cf(weight, g))
    res = x.movedim(1, -1)
    if output is not None:
        output.copy_(res.reshape(output.shape))
        res = output
    for _ in range(max(pad, 0)):
        res = res.squeeze(0)
    return res


def _finish(x, output, pad):
    res = x.movedim(1, -1)
    if output is not None:
        output.copy_(res.reshape(output.shape))
        res = output
    for _ in range(max(pad, 0)):
        res = res.squeeze(0)
    return res


def fmg_grid(hessian, gradient, weight=None, absolute=0, membrane=0, bending=0, lame=0, factor=1,
             voxel_size=1, bound="dft", nb_cycles=2, nb_iter=2, max_levels=16, solver="cg",
             output=None):
    """Full multigrid solve of (H + L) x = g (nitorch/_C/solve.py c_fmg_grid).

    solver: 'cg' (block-Jacobi PCG smoothing, the reference's default) or
    'relax' (multi-colour Gauss-Seidel).
    """
    if solver == "relaks":
        solver = "relax"
    if solver not in ("relax", "cg"):
        raise ValueError(f"fmg_grid: solver must be 'cg' or 'relax', got {solver!r}")
    _check(gradient, hessian)
    pad = 5 - gradient.dim()
    g = _to_cf(gradient, "gradient")
    h = _to_cf(hessian, "hessian")
    x = torch.zeros_like(g) if output is None else _to_cf(output, "output")
    fmg_grid_cf(h, g, x, _penalty(absolute, membrane, bending, lame, factor), voxel_size,
                bound, nb_cycles, nb_iter, max_levels, use_cg=(solver == "cg"), weight=_weight_cf(weight, g))
    return _finish(x, output, pad)


def pcg_grid(gradient, weight=None, hessian=None, absolute=0, membrane=0, bending=0, lame=0,
             factor=1, voxel_size=1, bound="dft", nb_iter=16, tol=1e-4, output=None):
    """Preconditioned conjugate gradient on (H + L) x = g
    (nitorch/_C/solve.py c_pcg_grid); `output` is the initial guess."""
    _check(gradient, hessian)
    pad = 5 - gradient.dim()
    g = _to_cf(gradient, "gradient")
    h = _to_cf(hessian, "hessian")
    x = torch.zeros_like(g) if output is None else _to_cf(output, "output")
    pcg_grid_cf(h, g, x, _penalty(absolute, membrane, bending, lame, factor), voxel_size, bound,
                nb_iter, tol, weight=_weight_cf(weight, g))
    return _finish(x, output, pad)


def precond_grid(gradient, weight=None, hessian=None, absolute=0, membrane=0, bending=0, lame=0,
                 factor=1, voxel_size=1, bound="dft", output=None):
    """Block-Jacobi preconditioner (H + diag L)^-1 g
    (nitorch/_C/solve.py c_precond_grid)."""
    _check(gradient, hessian)
    pad = 5 - gradient.dim()
    g = _to_cf(gradient, "gradient")
    h = _to_cf(hessian, "hessian")
    out = precond_grid_cf(h, g, _penalty(absolute, membrane, bending, lame, factor), voxel_size, bound,
                          weight=_weight_cf(weight, g))
    return _finish(out, output, pad)


def solve_grid_fmg(hessian, gradient, absolute=0, membrane=0, bending=0, lame=0, factor=1,
                   voxel_size=1, bound="dft", weights=None, optim="cg", nb_cycles=2, nb_iter=2):
    """nitorch.spatial.solve_grid_fmg (nitorch/spatial/_solvers.py:978-1039) on the GPU.
    Without spatial penalties the system is voxel-wise and solved in closed form."""
    lam = make_list(lame) if lame is not None else [0]
    if not (membrane or bending or any(lam)) and weights is None:
        return solve_grid_closedform(hessian, gradient, absolute=absolute, factor=factor,
                                     voxel_size=voxel_size)
    return fmg_grid(hessian, gradient, weights, absolute, membrane, bending, lame, factor, voxel_size,
                    bound, nb_cycles, nb_iter, solver=optim)


def solve_grid_closedform(hessian, gradient, absolute=0, factor=1, voxel_size=1):
    """(H + absolute * diag(v^2)) x = g voxel-wise (nitorch/spatial/_solvers.py
    solve_grid_closedform): small dense solves, done with torch on the GPU."""
    vs = make_list(voxel_size)
    vs = torch.as_tensor((vs + [vs[-1]] * 3)[:3], dtype=gradient.dtype, device=gradient.device)
    diag = absolute * factor * vs ** 2
    if hessian is None:
        return gradient / diag
    CC = hessian.shape[-1]
    if CC == 1:
        return gradient / (hessian + diag)
    if CC == 3:
        return gradient / (hessian + diag)
    H = torch.zeros(hessian.shape[:-1] + (3, 3), dtype=hessian.dtype, device=hessian.device)
    H[..., 0, 0], H[..., 1, 1], H[..., 2, 2] = hessian[..., 0], hessian[..., 1], hessian[..., 2]
    H[..., 0, 1] = H[..., 1, 0] = hessian[..., 3]
    H[..., 0, 2] = H[..., 2, 0] = hessian[..., 4]
    H[..., 1, 2] = H[..., 2, 1] = hessian[..., 5]
    H = H + torch.diag_embed(diag.expand(hessian.shape[:-1] + (3,)))
    return torch.linalg.solve(H, gradient.unsqueeze(-1)).squeeze(-1)


def solve_to_tolerance(hessian, gradient, penalty, voxel_size=(1, 1, 1), bound="dft",
                       tol=1e-4, nb_cycles=2, nb_iter=4, max_vcycles=20, max_levels=16, x=None,
                       workspace=None, use_cg=False, weight=None):
    """Plain GPU FMG followed by V-cycles on the finest level until
    ||g - (H + L) x|| / ||g|| <= tol.  Channel-first contiguous tensors.
    Returns (x, relative residual, number of extra V-cycles)."""
    g, h = gradient, hessian
    x = torch.zeros_like(g) if x is None else x
    CC = 0 if h is None else h.shape[1]
    dt = L.dtype_code(g.dtype)
    nbytes = _ws_fn()(dt, g.shape[0], L.i64arr(g.shape[2:]), CC, int(max_levels), int(bool(use_cg)),
                      int(weight is not None))
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=g.device)
    rel = ctypes.c_double(0)
    ncyc = ctypes.c_int32(0)
    L.check(L.lib().ni_fmg_solve(
        dt, g.shape[0], L.i64arr(g.shape[2:]), L.ptr(h), CC, L.ptr(weight), L.ptr(g), L.ptr(x), _bounds(bound),
        _voxel(voxel_size), _dvec(penalty), int(nb_cycles), int(nb_iter), int(max_levels),
        int(bool(use_cg)), float(tol), int(max_vcycles), L.ptr(workspace), int(nbytes),
        ctypes.byref(rel), ctypes.byref(ncyc), L.stream_of(g)), "fmg_solve")
    return x, rel.value, ncyc.value
