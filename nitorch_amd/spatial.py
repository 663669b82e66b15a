"""Public spatial API, drop-in for `nitorch.spatial.grid_pull / grid_push /
grid_count / grid_grad` (nitorch/spatial/_grid.py:90-328) plus
`identity_grid`, `add_identity_grid(_)`, `sub_identity_grid(_)`
(nitorch/spatial/_grid.py:443-560).

High-level shapes: input `(..., [channel], *inshape)`, grid
`(..., *outshape, dim)`; leading batch dimensions broadcast.  Computation
runs in the HIP backend (`nitorch_amd.grid`).
"""
import torch

from .grid import (GridPull, GridPush, GridCount, GridGrad, SplineCoeff, SplineCoeffND,
                   BoundType, InterpolationType, bound_to_nitorch, inter_to_nitorch)

__all__ = ["grid_pull", "grid_push", "grid_count", "grid_grad",
           "spline_coeff", "spline_coeff_nd",
           "identity_grid", "add_identity_grid", "add_identity_grid_",
           "sub_identity_grid", "sub_identity_grid_",
           "BoundType", "InterpolationType"]


def _expanded_shape(*shapes):
    """Broadcast shapes right-aligned (like torch.broadcast_shapes)."""
    return tuple(torch.broadcast_shapes(*[tuple(s) for s in shapes]))


def _preproc(grid, input=None, mode=None):
    """Broadcast / reshape to the backend's [B, C, *spatial] / [B, *spatial, D]
    (nitorch/spatial/_grid.py:90-127)."""
    dim = grid.shape[-1]
    if input is None:
        spatial = grid.shape[-dim - 1:-1]
        batch = grid.shape[:-dim - 1]
        grid = grid.reshape([-1, *spatial, dim])
        info = dict(batch=batch, channel=[1] if batch else [], dim=dim)
        return grid, info

    grid_spatial = grid.shape[-dim - 1:-1]
    grid_batch = grid.shape[:-dim - 1]
    input_spatial = input.shape[-dim:]
    channel = 0 if input.dim() == dim else input.shape[-dim - 1]
    input_batch = input.shape[:-dim - 1]

    if mode == "push":
        grid_spatial = input_spatial = _expanded_shape(grid_spatial, input_spatial)

    batch = _expanded_shape(grid_batch, input_batch)
    grid = grid.expand([*batch, *grid_spatial, dim])
    grid = grid.reshape([-1, *grid_spatial, dim])
    input = input.expand([*batch, channel or 1, *input_spatial])
    input = input.reshape([-1, channel or 1, *input_spatial])

    out_channel = [channel] if channel else ([1] if batch else [])
    info = dict(batch=batch, channel=out_channel, dim=dim)
    return grid, input, info


def _postproc(out, shape_info, mode):
    """nitorch/spatial/_grid.py:130-143"""
    dim = shape_info["dim"]
    if mode != "grad":
        spatial = out.shape[-dim:]
        feat = []
    else:
        spatial = out.shape[-dim - 1:-1]
        feat = [out.shape[-1]]
    batch = shape_info["batch"]
    channel = shape_info["channel"]
    return out.reshape([*batch, *channel, *spatial, *feat])


def grid_pull(input, grid, interpolation="linear", bound="zero",
              extrapolate=False, prefilter=False):
    """Sample an image with respect to a deformation field
    (nitorch/spatial/_grid.py:146-205).

    input : (..., [channel], *inshape) tensor
    grid  : (..., *outshape, dim) tensor of voxel coordinates
    returns (..., [channel], *outshape)

    Integer inputs are treated as label maps: each label is pulled as a soft
    mask and the arg-max label is returned.  With `prefilter`, the input is
    first turned into interpolating spline coefficients (`spline_coeff_nd`).
    For label maps this keeps the reference's statement
    `input = spline_coeff_nd(soft, ..., inplace=True)` (_grid.py:193-195):
    the prefiltered mask of one label replaces `input` for the next labels.
    """
    grid, input, shape_info = _preproc(grid, input)
    batch, channel = input.shape[:2]
    dim = grid.shape[-1]

    if not input.dtype.is_floating_point:
        out = input.new_zeros([batch, channel, *grid.shape[1:-1]])
        pmax = grid.new_zeros([batch, channel, *grid.shape[1:-1]])
        for label in input.unique():
            soft = (input == label).to(grid.dtype)
            if prefilter:
                input = spline_coeff_nd(soft, interpolation=interpolation, bound=bound, dim=dim,
                                        inplace=True)
            soft = GridPull.apply(soft, grid, interpolation, bound, extrapolate)
            out[soft > pmax] = label
            pmax = torch.max(pmax, soft)
    else:
        if prefilter:
            input = spline_coeff_nd(input, interpolation=interpolation, bound=bound, dim=dim)
        out = GridPull.apply(input, grid, interpolation, bound, extrapolate)
    return _postproc(out, shape_info, mode="pull")


def grid_push(input, grid, shape=None, interpolation="linear", bound="zero",
              extrapolate=False, prefilter=False):
    """Splat an image with respect to a deformation field (adjoint of pull,
    nitorch/spatial/_grid.py:208-251); with `prefilter`, the splatted image is
    turned into interpolating spline coefficients."""
    grid, input, shape_info = _preproc(grid, input, mode="push")
    dim = grid.shape[-1]
    if shape is None:
        shape = tuple(input.shape[2:])
    out = GridPush.apply(input, grid, shape, interpolation, bound, extrapolate)
    if prefilter:
        out = spline_coeff_nd(out, interpolation=interpolation, bound=bound, dim=dim, inplace=True)
    return _postproc(out, shape_info, mode="push")


def grid_count(grid, shape=None, interpolation="linear", bound="zero",
               extrapolate=False):
    """Splatting weights (push of ones, nitorch/spatial/_grid.py:254-287)."""
    grid, shape_info = _preproc(grid)
    if shape is None:
        shape = tuple(grid.shape[1:-1])
    out = GridCount.apply(grid, shape, interpolation, bound, extrapolate)
    return _postproc(out, shape_info, mode="count")


def grid_grad(input, grid, interpolation="linear", bound="zero",
              extrapolate=False, prefilter=False):
    """Spatial gradients of the spline at the grid points
    (nitorch/spatial/_grid.py:290-328); returns (..., [channel], *outshape, dim).
    With `prefilter`, the input is first turned into interpolating spline
    coefficients."""
    grid, input, shape_info = _preproc(grid, input)
    dim = grid.shape[-1]
    if prefilter:
        input = spline_coeff_nd(input, interpolation=interpolation, bound=bound, dim=dim)
    out = GridGrad.apply(input, grid, interpolation, bound, extrapolate)
    return _postproc(out, shape_info, mode="grad")


def spline_coeff(input, interpolation="linear", bound="dct2", dim=-1, inplace=False):
    """Interpolating spline coefficients along one dimension
    (nitorch/spatial/_grid.py:331-373): the B-spline coefficients whose
    order-`interpolation` spline passes through the samples, under the
    boundary condition `bound` (zero/dct1, replicate/dct2 or dft rules)."""
    return SplineCoeff.apply(input, bound, interpolation, dim, inplace)


def spline_coeff_nd(input, interpolation="linear", bound="dct2", dim=None, inplace=False):
    """Interpolating spline coefficients along the last `dim` dimensions
    (nitorch/spatial/_grid.py:376-421)."""
    return SplineCoeffND.apply(input, bound, interpolation, dim, inplace)


def identity_grid(shape, dtype=None, device=None, jitter=False):
    """Voxel-coordinate identity field of shape (*shape, dim)
    (nitorch/spatial/_grid.py:443-475: arange + meshgrid 'ij' + stack)."""
    mesh = torch.meshgrid(*[torch.arange(s, dtype=dtype, device=device) for s in shape],
                          indexing="ij")
    grid = torch.stack(mesh, dim=-1)
    if jitter:
        grid = grid + torch.rand_like(grid).sub_(0.5).mul_(jitter)
    return grid


def add_identity_grid_(disp):
    """In-place: displacement -> transformation (nitorch/spatial/_grid.py:478-500)."""
    dim = disp.shape[-1]
    spatial = disp.shape[-dim - 1:-1]
    mesh = torch.meshgrid(*[torch.arange(s, dtype=disp.dtype, device=disp.device) for s in spatial],
                          indexing="ij")
    for i, m in enumerate(mesh):
        disp[..., i].add_(m)
    return disp


def add_identity_grid(disp):
    return add_identity_grid_(disp.clone())


def sub_identity_grid_(grid):
    dim = grid.shape[-1]
    spatial = grid.shape[-dim - 1:-1]
    mesh = torch.meshgrid(*[torch.arange(s, dtype=grid.dtype, device=grid.device) for s in spatial],
                          indexing="ij")
    for i, m in enumerate(mesh):
        grid[..., i].sub_(m)
    return grid


def sub_identity_grid(grid):
    return sub_identity_grid_(grid.clone())
