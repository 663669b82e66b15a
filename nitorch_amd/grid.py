"""Backend module for spline push/pull on MI355X.

This module satisfies the plugin contract that `nitorch/_C/grid.py:9-79`
expects from a compiled backend: it exports

    grid_pull, grid_pull_backward, grid_push, grid_push_backward,
    grid_count, grid_count_backward, grid_grad, grid_grad_backward,
    InterpolationType, BoundType

with the reference's positional calling convention
`fn(input, grid, bound_list, interpolation_list, extrapolate_int)` (push and
count take `shape` after `grid`), and the autograd Functions
`GridPull / GridPush / GridCount / GridGrad` of `nitorch/_C/grid.py:192-331`.

Every call runs the HIP kernels of libnitorch_amd.so on torch's current
stream.  There is no CPU path: CPU tensors raise.
"""
import enum
import os

import torch

from . import _lib as L

COMPILED_BACKEND = "HIP"

if bool(int(os.environ.get("NI_SHOW_COMPILED_BACKEND", "0") or "0")):
    print("COMPILED_BACKEND:", COMPILED_BACKEND)


class BoundType(enum.IntEnum):
    """Boundary conditions (values of nitorch/_C/src/bounds.h:7-17, names and
    aliases of the pybind enum in nitorch/_C/spatial.cpp:48-62)."""
    replicate = 0
    nearest = 0
    dct1 = 1
    mirror = 1
    dct2 = 2
    reflect = 2
    dst1 = 3
    antimirror = 3
    dst2 = 4
    antireflect = 4
    dft = 5
    wrap = 5
    zero = 7


class InterpolationType(enum.IntEnum):
    """Spline orders (nitorch/_C/src/interpolation.h:17-19, spatial.cpp:64-73)."""
    nearest = 0
    linear = 1
    quadratic = 2
    cubic = 3
    fourth = 4
    fifth = 5
    sixth = 6
    seventh = 7


def make_list(x):
    if not isinstance(x, (list, tuple)):
        x = [x]
    return list(x)


def bound_to_nitorch(bound, as_type="str"):
    """Normalise boundary names (nitorch/_C/grid.py:92-138).

    as_type: 'str' -> names, 'enum' -> BoundType, 'int' -> enum values.
    """
    intype = type(bound)
    if not isinstance(bound, (list, tuple)):
        bound = [bound]
    obound = []
    for b in bound:
        if isinstance(b, BoundType):
            b = BoundType(b).name
        b = b.lower() if isinstance(b, str) else b
        if b in ("replicate", "repeat", "border", "nearest", 0):
            obound.append("replicate")
        elif b in ("zero", "zeros", "constant", 7):
            obound.append("zero")
        elif b in ("dct2", "reflect", "reflection", "neumann", 2):
            obound.append("dct2")
        elif b in ("dct1", "mirror", 1):
            obound.append("dct1")
        elif b in ("dft", "wrap", "circular", 5):
            obound.append("dft")
        elif b in ("dst2", "antireflect", "dirichlet", 4):
            obound.append("dst2")
        elif b in ("dst1", "antimirror", 3):
            obound.append("dst1")
        else:
            raise ValueError(f"Unknown boundary condition {b}")
    if as_type in ("enum", "int", int):
        obound = [getattr(BoundType, b) for b in obound]
        if as_type in ("int", int):
            obound = [int(b) for b in obound]
    if issubclass(intype, (list, tuple)):
        obound = intype(obound)
    else:
        obound = obound[0]
    return obound


def inter_to_nitorch(inter, as_type="str"):
    """Normalise interpolation orders (nitorch/_C/grid.py:141-186)."""
    intype = type(inter)
    if not isinstance(inter, (list, tuple)):
        inter = [inter]
    names = ["nearest", "linear", "quadratic", "cubic", "fourth", "fifth", "sixth", "seventh"]
    ointer = []
    for o in inter:
        if isinstance(o, InterpolationType):
            o = int(o)
        o = o.lower() if isinstance(o, str) else o
        if isinstance(o, str) and o in names:
            ointer.append(names.index(o))
        elif isinstance(o, int) and not isinstance(o, bool) and 0 <= o <= 7:
            ointer.append(o)
        else:
            raise ValueError(f"Unknown interpolation order {o}")
    if as_type in ("enum", "str", str):
        ointer = [InterpolationType(o) for o in ointer]
        if as_type in ("str", str):
            ointer = [o.name for o in ointer]
    if issubclass(intype, (list, tuple)):
        ointer = intype(ointer)
    else:
        ointer = ointer[0]
    return ointer


# --------------------------------------------------------------------------
# argument handling shared by the entry points (checks of pushpull.cpp:18-72)
# --------------------------------------------------------------------------
def _pad3(vals):
    vals = list(vals)
    if not vals:
        raise RuntimeError("(): expected non empty parameter list")
    return (vals + [vals[-1]] * 3)[:3]


def _codes(bound, interpolation):
    b = _pad3([int(bound_to_nitorch(x, "int")) for x in make_list(bound)])
    o = _pad3([int(inter_to_nitorch(x, "int")) for x in make_list(interpolation)])
    return L.i32arr(b), L.i32arr(o)


def _check(fname, input, grid, target_compat=False):
    if grid is None:
        raise RuntimeError(f"{fname}(): expected grid not be undefined")
    tensors = [grid] if input is None else [input, grid]
    for t in tensors:
        if t.layout != torch.strided:
            raise RuntimeError(f"{fname}(): expected tensors to have torch.strided layout")
        if t.dim() not in (3, 4, 5):
            raise RuntimeError(f"{fname}(): expected 3D, 4D or 5D tensors but got sizes {tuple(t.shape)}")
    dim = grid.dim() - 2
    if grid.shape[-1] != dim:
        raise RuntimeError(f"{fname}(): expected grid to have size {dim} in last dimension, "
                           f"but got grid with sizes {tuple(grid.shape)}")
    if input is not None:
        if input.device != grid.device:
            raise RuntimeError(f"{fname}(): expected input and grid to be on same device, but input is "
                               f"on {input.device} and grid is on {grid.device}")
        if input.dtype != grid.dtype:
            raise RuntimeError(f"{fname}(): expected input and grid to have the same dtype, but input has "
                               f"{input.dtype} and grid has {grid.dtype}")
        if input.dim() != dim + 2:
            raise RuntimeError(f"{fname}(): expected input with {dim + 2} dimensions, got {input.dim()}")
    for t in tensors:
        for s in t.shape[2:] if t is input else t.shape[1:-1]:
            if s <= 0:
                raise RuntimeError(f"{fname}(): expected tensors to have non-empty spatial dimensions, "
                                   f"but got sizes {tuple(t.shape)}")
    if target_compat:
        if input.shape[0] != grid.shape[0] or tuple(input.shape[2:]) != tuple(grid.shape[1:-1]):
            raise RuntimeError(f"{fname}(): expected grid and input to have same batch, width, height "
                               f"and (optionally) depth sizes, but got input with sizes {tuple(input.shape)} "
                               f"and grid with sizes {tuple(grid.shape)}")
    L.require_device(*tensors)
    L.dtype_code(grid.dtype)
    return dim


def _batch(src, N):
    """Broadcast a batch-1 tensor against N (stride-0 view); else require N."""
    if src.shape[0] == N:
        return src
    if src.shape[0] == 1:
        return src.expand((N,) + tuple(src.shape[1:]))
    raise RuntimeError(f"expected batch size {N} or 1, got {src.shape[0]}")


def _shape(shape, default):
    if shape is None or len(shape) == 0:
        return [int(s) for s in default]
    return [int(s) for s in shape]


# --------------------------------------------------------------------------
# backend entry points
# --------------------------------------------------------------------------
def _half_as_float(fn):
    """Half-precision tensors (the reference's GPU backend dispatches half,
    nitorch/_C/src/impl/pushpull_common.cpp:2505-2537, with float as its
    accumulation type) run the float32 kernels: float16 arguments are
    widened to float32 and the results narrowed back to float16."""
    import functools

    @functools.wraps(fn)
    def run(*args):
        if not any(torch.is_tensor(t) and t.dtype == torch.float16 for t in args):
            return fn(*args)
        wide = []
        for t in args:
            if torch.is_tensor(t) and t.dtype == torch.float16:
                w = t.float()
                w.requires_grad_(t.requires_grad)
                wide.append(w)
            else:
                wide.append(t)
        out = fn(*wide)
        if isinstance(out, list):
            return [o.half() for o in out]
        return out.half()
    return run


@_half_as_float
def grid_pull(input, grid, bound, interpolation, extrapolate):
    dim = _check("grid_pull", input, grid)
    N = grid.shape[0]
    input = _batch(input, N)
    C = input.shape[1]
    out = torch.empty((N, C) + tuple(grid.shape[1:-1]), dtype=input.dtype, device=input.device)
    b, o = _codes(bound, interpolation)
    L.check(L.lib().ni_grid_pull(
        L.dtype_code(input.dtype), dim, b, o, int(extrapolate), N, C,
        L.i64arr(input.shape[2:]), L.i64arr(grid.shape[1:-1]),
        L.ptr(input), L.strides(input), L.ptr(grid), L.strides(grid),
        L.ptr(out), L.strides(out), L.stream_of(grid)), "grid_pull")
    return out


@_half_as_float
def grid_grad(input, grid, bound, interpolation, extrapolate):
    dim = _check("grid_grad", input, grid)
    N = grid.shape[0]
    input = _batch(input, N)
    C = input.shape[1]
    out = torch.empty((N, C) + tuple(grid.shape[1:-1]) + (dim,), dtype=input.dtype, device=input.device)
    b, o = _codes(bound, interpolation)
    L.check(L.lib().ni_grid_grad(
        L.dtype_code(input.dtype), dim, b, o, int(extrapolate), N, C,
        L.i64arr(input.shape[2:]), L.i64arr(grid.shape[1:-1]),
        L.ptr(input), L.strides(input), L.ptr(grid), L.strides(grid),
        L.ptr(out), L.strides(out), L.stream_of(grid)), "grid_grad")
    return out


@_half_as_float
def grid_push(input, grid, shape, bound, interpolation, extrapolate):
    dim = _check("grid_push", input, grid, target_compat=True)
    N, C = input.shape[:2]
    shape = _shape(shape, input.shape[2:])
    if len(shape) != dim:
        raise RuntimeError(f"grid_push(): expected {dim} shape elements but got {len(shape)}")
    out = torch.empty([N, C] + shape, dtype=input.dtype, device=input.device)
    b, o = _codes(bound, interpolation)
    L.check(L.lib().ni_grid_push(
        L.dtype_code(input.dtype), dim, b, o, int(extrapolate), N, C,
        L.i64arr(shape), L.i64arr(grid.shape[1:-1]),
        L.ptr(input), L.strides(input), L.ptr(grid), L.strides(grid),
        L.ptr(out), L.strides(out), L.stream_of(grid)), "grid_push")
    return out


@_half_as_float
def grid_count(grid, shape, bound, interpolation, extrapolate):
    dim = _check("grid_count", None, grid)
    N = grid.shape[0]
    shape = _shape(shape, grid.shape[1:-1])
    if len(shape) != dim:
        raise RuntimeError(f"grid_count(): expected {dim} shape elements but got {len(shape)}")
    out = torch.empty([N, 1] + shape, dtype=grid.dtype, device=grid.device)
    b, o = _codes(bound, interpolation)
    L.check(L.lib().ni_grid_count(
        L.dtype_code(grid.dtype), dim, b, o, int(extrapolate), N,
        L.i64arr(shape), L.i64arr(grid.shape[1:-1]),
        L.ptr(grid), L.strides(grid), L.ptr(out), L.strides(out),
        L.stream_of(grid)), "grid_count")
    return out


@_half_as_float
def grid_pull_backward(grad, input, grid, bound, interpolation, extrapolate):
    """Returns [grad_input if input.requires_grad] + [grad_grid if grid.requires_grad]
    (the reference's deque, pushpull.cpp:109-128)."""
    dim = _check("grid_pull_backward", input, grid)
    N = grid.shape[0]
    input = _batch(input, N)
    C = input.shape[1]
    need_in, need_grid = input.requires_grad, grid.requires_grad
    out = []
    gi = torch.empty((N, C) + tuple(input.shape[2:]), dtype=input.dtype, device=input.device) if need_in else None
    gg = torch.empty(grid.shape, dtype=grid.dtype, device=grid.device) if need_grid else None
    if need_in or need_grid:
        grad = grad.detach()
        b, o = _codes(bound, interpolation)
        L.check(L.lib().ni_grid_pull_backward(
            L.dtype_code(input.dtype), dim, b, o, int(extrapolate), N, C,
            L.i64arr(input.shape[2:]), L.i64arr(grid.shape[1:-1]),
            L.ptr(grad), L.strides(grad), L.ptr(input), L.strides(input),
            L.ptr(grid), L.strides(grid), L.ptr(gi), L.strides(gi),
            L.ptr(gg), L.strides(gg), L.stream_of(grid)), "grid_pull_backward")
    if need_in:
        out.append(gi)
    if need_grid:
        out.append(gg)
    return out


@_half_as_float
def grid_push_backward(grad, input, grid, bound, interpolation, extrapolate):
    dim = _check("grid_push_backward", input, grid)
    N, C = input.shape[:2]
    need_in, need_grid = input.requires_grad, grid.requires_grad
    out = []
    gi = torch.empty(input.shape, dtype=input.dtype, device=input.device) if need_in else None
    gg = torch.empty(grid.shape, dtype=grid.dtype, device=grid.device) if need_grid else None
    if need_in or need_grid:
        grad = grad.detach()
        b, o = _codes(bound, interpolation)
        L.check(L.lib().ni_grid_push_backward(
            L.dtype_code(input.dtype), dim, b, o, int(extrapolate), N, C,
            L.i64arr(grad.shape[2:]), L.i64arr(grid.shape[1:-1]),
            L.ptr(grad), L.strides(grad), L.ptr(input), L.strides(input),
            L.ptr(grid), L.strides(grid), L.ptr(gi), L.strides(gi),
            L.ptr(gg), L.strides(gg), L.stream_of(grid)), "grid_push_backward")
    if need_in:
        out.append(gi)
    if need_grid:
        out.append(gg)
    return out


@_half_as_float
def grid_count_backward(grad, grid, bound, interpolation, extrapolate):
    dim = _check("grid_count_backward", None, grid)
    N = grid.shape[0]
    gg = torch.empty(grid.shape, dtype=grid.dtype, device=grid.device)
    grad = _batch(grad.detach(), N)
    b, o = _codes(bound, interpolation)
    L.check(L.lib().ni_grid_count_backward(
        L.dtype_code(grid.dtype), dim, b, o, int(extrapolate), N,
        L.i64arr(grad.shape[2:]), L.i64arr(grid.shape[1:-1]),
        L.ptr(grad), L.strides(grad), L.ptr(grid), L.strides(grid),
        L.ptr(gg), L.strides(gg), L.stream_of(grid)), "grid_count_backward")
    return gg


@_half_as_float
def grid_grad_backward(grad, input, grid, bound, interpolation, extrapolate):
    dim = _check("grid_grad_backward", input, grid)
    N = grid.shape[0]
    input = _batch(input, N)
    C = input.shape[1]
    need_in, need_grid = input.requires_grad, grid.requires_grad
    out = []
    gi = torch.empty((N, C) + tuple(input.shape[2:]), dtype=input.dtype, device=input.device) if need_in else None
    gg = torch.empty(grid.shape, dtype=grid.dtype, device=grid.device) if need_grid else None
    if need_in or need_grid:
        grad = grad.detach()
        b, o = _codes(bound, interpolation)
        L.check(L.lib().ni_grid_grad_backward(
            L.dtype_code(input.dtype), dim, b, o, int(extrapolate), N, C,
            L.i64arr(input.shape[2:]), L.i64arr(grid.shape[1:-1]),
            L.ptr(grad), L.strides(grad), L.ptr(input), L.strides(input),
            L.ptr(grid), L.strides(grid), L.ptr(gi), L.strides(gi),
            L.ptr(gg), L.strides(gg), L.stream_of(grid)), "grid_grad_backward")
    if need_in:
        out.append(gi)
    if need_grid:
        out.append(gg)
    return out


# --------------------------------------------------------------------------
# autograd Functions (nitorch/_C/grid.py:192-331)
# --------------------------------------------------------------------------
def _opts(bound, interpolation, extrapolate):
    bound = bound_to_nitorch(make_list(bound), as_type="enum")
    interpolation = inter_to_nitorch(make_list(interpolation), as_type="enum")
    return bound, interpolation, int(extrapolate)


def _split(ctx, grads, first):
    """Map the backend's list to (grad_first, grad_grid) like the reference."""
    g_first = g_grid = None
    if ctx.needs_input_grad[first]:
        g_first = grads[0]
        if ctx.needs_input_grad[first + 1]:
            g_grid = grads[1]
    elif ctx.needs_input_grad[first + 1]:
        g_grid = grads[0]
    return g_first, g_grid


def _with_req(t, req):
    t = t.detach()
    t.requires_grad_(bool(req))
    return t


class GridPull(torch.autograd.Function):

    @staticmethod
    def forward(ctx, input, grid, interpolation, bound, extrapolate):
        opt = _opts(bound, interpolation, extrapolate)
        output = grid_pull(input, grid, *opt)
        ctx.opt = opt
        ctx.save_for_backward(input, grid)
        return output

    @staticmethod
    def backward(ctx, grad):
        input, grid = ctx.saved_tensors
        need_in, need_grid = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_in or need_grid):
            return None, None, None, None, None
        grads = grid_pull_backward(grad, _with_req(input, need_in), _with_req(grid, need_grid), *ctx.opt)
        gi, gg = _split(ctx, grads, 0)
        if gi is not None and gi.shape[0] != input.shape[0]:
            gi = gi.sum(0, keepdim=True)      # batch-1 input broadcast against the grid
        return gi, gg, None, None, None


class GridPush(torch.autograd.Function):

    @staticmethod
    def forward(ctx, input, grid, shape, interpolation, bound, extrapolate):
        opt = _opts(bound, interpolation, extrapolate)
        output = grid_push(input, grid, shape, *opt)
        ctx.opt = opt
        ctx.save_for_backward(input, grid)
        return output

    @staticmethod
    def backward(ctx, grad):
        input, grid = ctx.saved_tensors
        need_in, need_grid = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_in or need_grid):
            return None, None, None, None, None, None
        grads = grid_push_backward(grad, _with_req(input, need_in), _with_req(grid, need_grid), *ctx.opt)
        gi, gg = _split(ctx, grads, 0)
        return gi, gg, None, None, None, None


class GridCount(torch.autograd.Function):

    @staticmethod
    def forward(ctx, grid, shape, interpolation, bound, extrapolate):
        opt = _opts(bound, interpolation, extrapolate)
        output = grid_count(grid, shape, *opt)
        ctx.opt = opt
        ctx.save_for_backward(grid)
        return output

    @staticmethod
    def backward(ctx, grad):
        (grid,) = ctx.saved_tensors
        grad_grid = None
        if ctx.needs_input_grad[0]:
            grad_grid = grid_count_backward(grad, grid, *ctx.opt)
        return grad_grid, None, None, None, None


class GridGrad(torch.autograd.Function):

    @staticmethod
    def forward(ctx, input, grid, interpolation, bound, extrapolate):
        opt = _opts(bound, interpolation, extrapolate)
        output = grid_grad(input, grid, *opt)
        ctx.opt = opt
        ctx.save_for_backward(input, grid)
        return output

    @staticmethod
    def backward(ctx, grad):
        input, grid = ctx.saved_tensors
        need_in, need_grid = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_in or need_grid):
            return None, None, None, None, None
        grads = grid_grad_backward(grad, _with_req(input, need_in), _with_req(grid, need_grid), *ctx.opt)
        gi, gg = _split(ctx, grads, 0)
        if gi is not None and gi.shape[0] != input.shape[0]:
            gi = gi.sum(0, keepdim=True)
        return gi, gg, None, None, None


# --------------------------------------------------------------------------
# Interpolating spline coefficients (prefilter).  Same entry points as the
# ones nitorch/_C/grid.py:4 imports from interpol.coeff (restated in the
# reference's nitorch/_C/_ts/coeff.py:278-344) and the autograd Functions
# SplineCoeff / SplineCoeffND of nitorch/_C/grid.py:335-390.
# --------------------------------------------------------------------------
_COEFF_BOUNDS = (BoundType.zero, BoundType.dct1, BoundType.replicate, BoundType.dct2, BoundType.dft)


def _coeff_run(x, nbatch, shape, bound, order):
    """In-place prefilter of the contiguous tensor x = [nbatch, *shape] along
    every axis of `shape` (bound / order: one entry per axis)."""
    for n, b, o in zip(shape, bound, order):
        if o >= 2 and n > 1 and b not in _COEFF_BOUNDS:
            raise NotImplementedError(
                f"spline_coeff: boundary condition {BoundType(b).name} is not supported "
                "(zero, dct1, replicate, dct2 and dft are)")
    if x.numel() == 0 or all(o < 2 for o in order):
        return x
    L.require_device(x)
    rc = L.lib().ni_spline_coeff(L.dtype_code(x.dtype), int(nbatch), len(shape), L.i64arr(shape),
                                 L.i32arr(bound), L.i32arr(order), L.ptr(x), L.stream_of(x))
    L.check(rc, "spline_coeff")
    return x


def _coeff_target(inp, inplace):
    """(contiguous working tensor, whether to copy it back into `inp`)."""
    if inplace and inp.is_contiguous():
        return inp, False
    return inp.clone(memory_format=torch.contiguous_format), inplace


def spline_coeff(inp, bound, order, dim=-1, inplace=False):
    """Interpolating coefficients along one dimension (`interpol.coeff.
    spline_coeff(inp, bound, order, dim, inplace)`, nitorch/_C/_ts/coeff.py:
    278-306); bound and order are single codes (BoundType / 0..7)."""
    bound = int(bound_to_nitorch(bound[0] if isinstance(bound, (list, tuple)) else bound, "int"))
    order = int(inter_to_nitorch(order[0] if isinstance(order, (list, tuple)) else order, "int"))
    if inp.dim() == 0:
        return inp if inplace else inp.clone()
    dim = dim % inp.dim()
    x, copy_back = _coeff_target(inp, inplace)
    outer = 1
    for s in x.shape[:dim]:
        outer *= s
    inner = 1
    for s in x.shape[dim + 1:]:
        inner *= s
    _coeff_run(x, outer, [x.shape[dim], inner], [bound, bound], [order, 0])
    if copy_back:
        inp.copy_(x)
        return inp
    return x


def spline_coeff_nd(inp, bound, order, dim=None, inplace=False):
    """Interpolating coefficients along the last `dim` dimensions, first to
    last (`interpol.coeff.spline_coeff_nd`, nitorch/_C/_ts/coeff.py:309-344);
    bound / order lists are padded by repeating their last entry."""
    dim = inp.dim() if dim is None else int(dim)
    bound = [int(bound_to_nitorch(b, "int")) for b in make_list(bound)]
    order = [int(inter_to_nitorch(o, "int")) for o in make_list(order)]
    bound = (bound + [bound[-1]] * dim)[:dim]
    order = (order + [order[-1]] * dim)[:dim]
    if dim == 0:
        return inp if inplace else inp.clone()
    if dim > 8:
        raise ValueError(f"spline_coeff_nd: at most 8 dimensions are supported (got {dim})")
    x, copy_back = _coeff_target(inp, inplace)
    nbatch = 1
    for s in x.shape[:x.dim() - dim]:
        nbatch *= s
    _coeff_run(x, nbatch, list(x.shape[x.dim() - dim:]), bound, order)
    if copy_back:
        inp.copy_(x)
        return inp
    return x


class SplineCoeff(torch.autograd.Function):

    @staticmethod
    def forward(ctx, input, bound, interpolation, dim, inplace):
        opt = (bound_to_nitorch(bound, as_type="int"), inter_to_nitorch(interpolation, as_type="int"),
               dim, inplace)
        output = spline_coeff(input, *opt)
        if input.requires_grad:
            ctx.opt = opt
        return output

    @staticmethod
    def backward(ctx, grad):
        # the reference treats the filter as symmetric: backward == forward
        grad = spline_coeff(grad, *ctx.opt[:-1], inplace=False)
        return grad, None, None, None, None


class SplineCoeffND(torch.autograd.Function):

    @staticmethod
    def forward(ctx, input, bound, interpolation, dim, inplace):
        opt = (bound_to_nitorch(make_list(bound), as_type="int"),
               inter_to_nitorch(make_list(interpolation), as_type="int"), dim, inplace)
        output = spline_coeff_nd(input, *opt)
        if input.requires_grad:
            ctx.opt = opt
        return output

    @staticmethod
    def backward(ctx, grad):
        grad = spline_coeff_nd(grad, *ctx.opt[:-1], inplace=False)
        return grad, None, None, None, None
