"""Compiled resize, prolongation and restriction on the HIP backend.

* the resize bindings of nitorch/_C/spatial.cpp:99-108 -- `resize`,
  `resize_backward`, `prolongation`, `prolongation_backward`,
  `restriction`, `restriction_backward` (nitorch/_C/src/resize.cpp:25-124),
  on `ni_resize`;
* the Python layer nitorch/_C/resize.py:1-232 (`align_to_nitorch`,
  `c_resize`, `c_prolongation`, `c_restriction`, `c_prolongation_grid`,
  `c_restriction_grid`); as in the reference, `c_prolongation` /
  `c_restriction` run the FMG transfers (`fmg_prolongation` /
  `fmg_restriction`, nitorch/_C/src/multires.cpp:24-76).

Every call runs libnitorch_amd.so on torch's current stream; CPU tensors
raise.  Image-like tensors are [N, C, *spatial] with 1 to 3 spatial dims.
"""
import enum
import math

import torch

from . import _lib as L
from .grid import bound_to_nitorch, inter_to_nitorch

L.register_signatures({
    "ni_resize": [L._I32, L._I32, L._I64, L._I64, L._PI64, L._PI64, L.ctypes.POINTER(L.ctypes.c_double),
                  L._PI32, L._PI32, L._PI32, L._I32, L._I32, L._I32, L._P, L._P, L._P],
})


class GridAlignType(enum.IntEnum):
    """nitorch/_C/src/grid_align.h:7-12 (pybind names, spatial.cpp:75-79)."""
    edge = 0
    center = 1
    first = 2
    last = 3


def make_list(x, n=None):
    """A value or a sequence as a list, extended to `n` entries by repeating
    its last one."""
    out = list(x) if isinstance(x, (list, tuple)) else [x]
    if n and len(out) < n:
        out += out[-1:] * (n - len(out))
    return out


_ALIGN_NAMES = {"c": "center", "e": "edge", "f": "first", "l": "last"}


def align_to_nitorch(align, as_type="str"):
    """Normalise grid-alignment modes (nitorch/_C/resize.py:16-53): each mode
    is a `GridAlignType` or a string whose first letter is c / e / f / l
    ('centers', 'edges', 'first', 'last').  Returns names, or enum members
    when `as_type` is 'enum' / 'int' / int; a list or tuple in gives the same
    container out, a single mode a single value."""
    def name_of(mode):
        if isinstance(mode, GridAlignType):
            return mode.name
        key = mode.lower()[:1] if isinstance(mode, str) else None
        if key not in _ALIGN_NAMES:
            raise ValueError(f"Unknown alignment mode {mode}")
        return _ALIGN_NAMES[key]

    as_enum = as_type in ("enum", "int", int)
    convert = (lambda m: GridAlignType[name_of(m)]) if as_enum else name_of
    if isinstance(align, (list, tuple)):
        return type(align)(convert(m) for m in align)
    return convert(align)


def _pad3(vals):
    vals = list(vals)
    return (vals + [vals[-1]] * 3)[:3]


def _resize(input, output, factor, bound, interpolation, mode, do_adjoint, normalize):
    """ni::resize (resize.cpp:25-59): `output` may be None (allocated with
    the reference's prepare_tensors shapes); a given output is added into by
    the generic kernels and overwritten by the isotropic nearest / linear
    (and 3-D quadratic) ones, or zeroed first for the adjoint."""
    if input is None or input.numel() == 0:
        raise RuntimeError("resize(): expected a non empty input")
    dim = input.dim() - 2
    if dim < 1 or dim > 3:
        raise RuntimeError("resize(): expected a 1D, 2D or 3D tensor")
    L.require_device(input)
    factor = [float(f) for f in (factor if len(factor) else [1.])]
    f3 = [factor[0], factor[1] if len(factor) > 1 else factor[0], 0.]
    f3[2] = factor[2] if len(factor) > 2 else f3[1]
    bound = _pad3([int(bound_to_nitorch(b, "int")) for b in (bound or ["dct2"])])
    order = _pad3([int(inter_to_nitorch(o, "int")) for o in (interpolation or [1])])
    mode = _pad3([int(align_to_nitorch(m, "int")) for m in (mode or ["center"])])
    given = output is not None and output.numel() > 0
    if not given:
        # prepare_tensors (resize_common.cpp:1490-1545)
        if do_adjoint:
            shape = [int(math.ceil(s / f)) for s, f in zip(input.shape[2:], f3)]
        else:
            shape = [int(math.floor(s * f)) for s, f in zip(input.shape[2:], f3)]
        output = input.new_zeros([*input.shape[:2], *shape])
    else:
        if output.shape[:2] != input.shape[:2] or output.dim() != input.dim():
            raise RuntimeError("resize(): expected input and output with the same batch and channel sizes")
        if output.dtype != input.dtype or output.device != input.device:
            raise RuntimeError("resize(): expected input and output on the same device and of the same dtype")
    if output.numel() == 0:
        raise RuntimeError("resize(): expected a non empty output")
    x = input.contiguous()
    out = output if output.is_contiguous() else output.contiguous()
    rc = L.lib().ni_resize(L.dtype_code(x.dtype), dim, x.shape[0], x.shape[1], L.i64arr(x.shape[2:]),
                           L.i64arr(out.shape[2:]), (L.ctypes.c_double * 3)(*f3), L.i32arr(bound),
                           L.i32arr(order), L.i32arr(mode), int(bool(do_adjoint)), int(bool(normalize)),
                           int(given), L.ptr(x), L.ptr(out), L.stream_of(x))
    L.check(rc, "resize")
    if out is not output:
        output.copy_(out)
    return output


def resize(input, output=None, factor=(), bound=(), interpolation=(), mode=(), do_adjoint=False,
           normalize=False):
    """ni::resize (resize.cpp:25-59); defaults: factor 1, dct2, linear, center."""
    return _resize(input, output, list(factor), list(bound) or ["dct2"], list(interpolation) or [1],
                   list(mode) or ["center"], do_adjoint, normalize)


def resize_backward(grad, output=None, factor=(), bound=(), interpolation=(), mode=(), do_adjoint=False,
                    normalize=False):
    """ni::resize_backward (resize.cpp:89-102): resize with inverted factors
    and the adjoint flag flipped."""
    ifactor = [1. / f for f in factor]
    return resize(grad, output, ifactor, bound, interpolation, mode, not do_adjoint, normalize)


def prolongation(input, output=None, bound=(), interpolation=()):
    """ni::prolongation (resize.cpp:61-72): quadratic, edge-aligned."""
    return _resize(input, output, [1.], list(bound) or ["dct2"], list(interpolation) or [2],
                   [GridAlignType.edge], False, False)


def restriction(input, output=None, bound=(), interpolation=()):
    """ni::restriction (resize.cpp:74-85): linear adjoint, edge-aligned,
    normalised."""
    return _resize(input, output, [1.], list(bound) or ["dct2"], list(interpolation) or [1],
                   [GridAlignType.edge], True, True)


def prolongation_backward(grad, output=None, bound=(), interpolation=()):
    """ni::prolongation_backward (resize.cpp:104-113)."""
    return _resize(grad, output, [1.], list(bound) or ["dct2"], list(interpolation) or [2],
                   [GridAlignType.edge], True, False)


def restriction_backward(grad, output=None, bound=(), interpolation=()):
    """ni::restriction_backward (resize.cpp:115-124)."""
    return _resize(grad, output, [1.], list(bound) or ["dct2"], list(interpolation) or [1],
                   [GridAlignType.edge], False, True)


def c_resize(input, factor=None, bound="dct2", interpolation=1, mode="center",
             shape=None, output=None, adjoint=False, normalize=False):
    """Resize a spatial tensor (nitorch/_C/resize.py:64-97).

    input : (N, C, *inshape) tensor.  `factor`, `bound`, `interpolation`
    and `mode` are values or per-dimension sequences (the last entry
    repeats).  With `shape` (and no `output`), the result is allocated at
    that shape; otherwise its shape follows from the factor.  At least one
    of `factor` and `shape` is required.  Returns (N, C, *shape)."""
    if torch.is_tensor(factor):
        factor = factor.double().tolist()
    if not factor and not shape:
        raise ValueError("c_resize: give a factor, a shape or both")
    if output is None and shape:
        output = input.new_zeros(tuple(input.shape[:2]) + tuple(shape))
    return _resize(input, output, make_list(factor or 1.), bound_to_nitorch(make_list(bound), "int"),
                   inter_to_nitorch(make_list(interpolation), "int"), align_to_nitorch(make_list(mode), "int"),
                   adjoint, normalize)


def c_prolongation(input, factor=2., bound="dct2", shape=None, output=None):
    """FMG prolongation into a (N, C, *shape) output (nitorch/_C/resize.py:99-124)."""
    from . import solve as SV
    bound = bound_to_nitorch(make_list(bound), "int")
    if output is None:
        if not shape:
            factor = make_list(factor, 3)
            shape = [int((s * f) // 1) for s, f in zip(input.shape[2:], factor)]
        output = input.new_zeros([*input.shape[:2], *shape])
    return SV.fmg_prolongation(input, output, bound)


def c_restriction(input, factor=2., bound="dct2", shape=None, output=None):
    """FMG restriction into a (N, C, *shape) output (nitorch/_C/resize.py:126-151)."""
    from . import solve as SV
    bound = bound_to_nitorch(make_list(bound), "int")
    if output is None:
        if not shape:
            factor = make_list(factor, 3)
            shape = [int(math.ceil(s / f)) for s, f in zip(input.shape[2:], factor)]
        output = input.new_zeros([*input.shape[:2], *shape])
    return SV.fmg_restriction(input, output, bound)


def _rescale_components(output, factor, dim, hessian):
    if not hessian:
        for c in range(dim):
            output[..., c] *= factor[c]
    else:
        count = dim
        for c in range(dim):
            output[..., c] *= factor[c] ** 2
            for cc in range(c + 1, dim):
                output[..., count] *= factor[c] * factor[cc]
                count += 1
    return output


def c_prolongation_grid(input, factor=2., bound="dct2", shape=None, hessian=False, output=None):
    """Prolongation of a displacement (or Hessian) field (N, *inshape, D)
    with per-component rescaling (nitorch/_C/resize.py:153-191)."""
    dim = input.shape[-1]
    input = input.movedim(-1, 1)
    if output is not None:
        output = output.movedim(-1, 1)
    output = c_prolongation(input, factor, bound, shape, output)
    output = output.movedim(1, -1)
    # as the reference: the trailing `dim` sizes of the channel-last output
    factor = [o / i for i, o in zip(input.shape[-dim:], output.shape[-dim:])]
    return _rescale_components(output, factor, dim, hessian)


def c_restriction_grid(input, factor=2., bound="dct2", shape=None, hessian=False, output=None):
    """Restriction of a displacement (or Hessian) field (N, *inshape, D)
    with per-component rescaling (nitorch/_C/resize.py:193-232)."""
    dim = input.shape[-1]
    input = input.movedim(-1, 1)
    if output is not None:
        output = output.movedim(-1, 1)
    output = c_restriction(input, factor, bound, shape, output)
    output = output.movedim(1, -1)
    # as the reference: the trailing `dim` sizes of the channel-last output
    factor = [o / i for i, o in zip(input.shape[-dim:], output.shape[-dim:])]
    return _rescale_components(output, factor, dim, hessian)
