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
           "spline_coeff", "spline_coeff_nd", "affine_grid", "resize", "resize_grid", "reslice",
           "grid_jacobian",
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


# --------------------------------------------------------------------------
# Resampling on affine lattices (nitorch/spatial/_grid.py:579-915): grids
# built with torch on the device, sampled by grid_pull (prefiltered by
# default, as in the reference).
# --------------------------------------------------------------------------
def _matvec(mat, vec):
    # broadcast multiply-and-sum: one elementwise kernel, not a batched GEMM
    # over every voxel
    return (mat * vec[..., None, :]).sum(-1)


def _pad_last(vals, n):
    vals = list(vals)
    return (vals + vals[-1:] * n)[:n]


def affine_grid(mat, shape, jitter=False):
    """Grid of the voxel coordinates mat @ [i, j, k, 1] of a lattice of
    `shape` (nitorch/spatial/_grid.py:579-615): (*batch, *shape, dim)."""
    mat = torch.as_tensor(mat)
    shape = list(shape)
    nb_dim = mat.shape[-1] - 1
    if nb_dim != len(shape):
        raise ValueError("Dimension of the affine matrix ({}) and shape ({}) "
                         "are not the same.".format(nb_dim, len(shape)))
    if mat.shape[-2] not in (nb_dim, nb_dim + 1):
        raise ValueError("First argument should be matrces of shape (..., {0}, {1}) or "
                         "(..., {1}, {1}) but got {2}.".format(nb_dim, nb_dim + 1, mat.shape))
    batch_shape = mat.shape[:-2]
    grid = identity_grid(shape, mat.dtype, mat.device, jitter=jitter)
    if batch_shape:
        grid = grid.reshape([1] * len(batch_shape) + list(grid.shape))
        mat = mat.reshape(list(batch_shape) + [1] * nb_dim + list(mat.shape[-2:]))
    return _matvec(mat[..., :nb_dim, :nb_dim], grid) + mat[..., :nb_dim, -1]


def resize(image, factor=None, shape=None, affine=None, anchor="c", dim=None, *args, **kwargs):
    """Resize an image by a factor or to a shape (nitorch/spatial/_grid.py:618-754).

    anchor per dimension: 'c' (corner centres), 'e' (outer edges), 'f'
    (first voxel), 'l' (last voxel); extra arguments go to grid_pull, with
    prefilter=True by default.  Returns the resized image and, with
    `affine`, its orientation matrix."""
    from .affine import affine_resize
    image = torch.as_tensor(image)
    nb_dim = dim or (len(shape) if shape is not None else image.dim() - 2)
    inshape = image.shape[-nb_dim:]
    info = {"dtype": image.dtype, "device": image.device}
    if torch.is_tensor(factor):
        factor = factor.tolist()
    factor = _pad_last(factor if isinstance(factor, (list, tuple)) else [factor or 0.], nb_dim)
    factor = [float(f or 0.) for f in factor]
    outshape = _pad_last(shape if shape is not None else [None], nb_dim)
    anchor = [a[0].lower() for a in _pad_last(anchor if isinstance(anchor, (list, tuple)) else [anchor],
                                               nb_dim)]
    return_trf = kwargs.pop("_return_trf", False)

    outshape = [o or max(1, int(i * f)) for i, o, f in zip(inshape, outshape, factor)]
    if any(not s for s in outshape):
        raise ValueError("Either factor or shape must be set in all dimensions")
    factor = [f or o / i for o, i, f in zip(outshape, inshape, factor)]

    # input coordinate = scale * output coordinate + shift, per dimension
    lin, scales, shifts = [], [], []
    for anch, f, inshp, outshp in zip(anchor, factor, inshape, outshape):
        if inshp == 1 or outshp == 1:
            anch = "e"
        if anch == "c":
            lin.append(torch.linspace(0, inshp - 1, outshp, **info))
            scales.append((inshp - 1) / (outshp - 1))
            shifts.append(0)
        elif anch == "e":
            scale = inshp / outshp
            shift = 0.5 * (scale - 1)
            lin.append(torch.arange(0., outshp, **info) * scale + shift)
            scales.append(scale)
            shifts.append(shift)
        elif anch == "f":
            lin.append(torch.arange(0., outshp, **info) / f)
            scales.append(1 / f)
            shifts.append(0)
        elif anch == "l":
            shift = (inshp - 1) - (outshp - 1) / f
            lin.append(torch.arange(0., outshp, **info) / f + shift)
            scales.append(1 / f)
            shifts.append(shift)
        else:
            raise ValueError("Unknown anchor {}".format(anch))
        if inshp == outshp == 1:
            scales[-1] = 1 / f
    grid = torch.stack(torch.meshgrid(*lin, indexing="ij"), dim=-1)

    kwargs.setdefault("prefilter", True)
    resized = grid_pull(image, grid, *args, **kwargs)

    if affine is not None:
        affine, _ = affine_resize(affine, inshape, factor, anchor, shape_out=outshape)
        if return_trf:
            return resized, affine, (scales, shifts)
        return resized, affine
    if return_trf:
        return resized, (scales, shifts)
    return resized


def resize_grid(grid, factor=None, shape=None, type="grid", affine=None, *args, **kwargs):
    """Resize a transformation (type 'grid') or displacement field and
    express it in the resized voxel space (nitorch/spatial/_grid.py:757-858)."""
    kwargs["_return_trf"] = True
    kwargs.setdefault("prefilter", True)
    ndim = grid.shape[-1]
    spatial = grid.shape[-ndim - 1:-1]
    batch = grid.shape[:-ndim - 1]
    if len(batch) == 0:
        grid = grid[None]
    elif len(batch) > 1:
        grid = grid.reshape([-1, *spatial, ndim])
    if affine is not None:
        affine = affine.expand([*batch, *affine.shape[-2:]])
        affine = affine.reshape([-1, *affine.shape[-2:]])
    grid = grid.movedim(-1, 1)
    outputs = resize(grid, factor, shape, affine, *args, **kwargs)
    if affine is not None:
        grid, affine, (scales, shifts) = outputs
    else:
        grid, (scales, shifts) = outputs
    grid = grid.movedim(1, -1)
    # original = scale * resized + shift  ->  resized = (original - shift) / scale
    comps = []
    for d, (scl, shft) in enumerate(zip(scales, shifts)):
        g1 = grid[..., d]
        if type[0].lower() == "g":
            g1 = g1 - shft
        comps.append(g1 / scl)
    grid = torch.stack(comps, -1)
    grid = grid.reshape([*batch, *grid.shape[-ndim - 1:]])
    if affine is not None:
        affine = affine.reshape([*batch, *affine.shape[-2:]])
        return grid, affine
    return grid


def reslice(image, affine, affine_to, shape_to=None, **kwargs):
    """Resample an image into another space (shape + affine)
    (nitorch/spatial/_grid.py:861-915); grid_pull options, prefilter=True
    by default."""
    from .affine import affine_lmdiv
    kwargs.setdefault("prefilter", True)
    image = torch.as_tensor(image)
    backend = dict(dtype=image.dtype, device=image.device)
    if not backend["dtype"].is_floating_point:
        backend["dtype"] = torch.get_default_dtype()
    affine = torch.as_tensor(affine).to(**backend)
    affine_to = torch.as_tensor(affine_to).to(**backend)
    dim = affine.shape[-1] - 1
    batch = affine.shape[:-2]
    channels = image.shape[len(batch):-dim]
    shape = image.shape[-dim:]
    if shape_to is None:
        shape_to = shape

    def prod(s):
        p = 1
        for v in s:
            p *= v
        return p

    transformation = affine_lmdiv(affine, affine_to)
    grid = affine_grid(transformation, shape_to)
    if not prod(batch):
        grid = grid[None]
    image = image.reshape([prod(batch) or 1, prod(channels) or 1, *shape])
    image = grid_pull(image, grid, **kwargs)
    return image.reshape([*batch, *channels, *shape_to])


def _diff_central(x, dims, bound="dct2", voxel_size=1):
    """First-order central differences of x along `dims`, stacked last
    (nitorch/spatial/_finite_differences.py:154-184, `diff(side='c')`):
    (x[i+1] - x[i-1]) / (2 vs) inside, the bound's rule at both ends, and a
    copy of x along an axis of length 1.  Each axis's differences are
    computed into a dense tensor of x's layout (no intermediate copies)."""
    vs = list(voxel_size) if isinstance(voxel_size, (list, tuple)) else [voxel_size]
    vs = (vs + vs[-1:] * len(dims))[:len(dims)]
    if bound not in ("dct2", "replicate", "dct1", "dst2", "dst1", "zero", "dft"):
        raise ValueError("Unknown boundary type {}.".format(bound))
    out = []
    for d, v in zip(dims, vs):
        xm = x.movedim(d, 0)
        if xm.shape[0] == 1:
            out.append(x.clone())
            continue
        df = torch.empty_like(xm)
        torch.sub(xm[2:], xm[:-2], out=df[1:-1])
        if bound in ("dct2", "replicate"):
            torch.sub(xm[1], xm[0], out=df[0])
            torch.sub(xm[-1], xm[-2], out=df[-1])
        elif bound == "dct1":
            df[0].zero_()
            df[-1].zero_()
        elif bound == "dst2":
            torch.add(xm[1], xm[0], out=df[0])
            torch.add(xm[-1], xm[-2], out=df[-1]).neg_()
        elif bound in ("dst1", "zero"):
            df[0].copy_(xm[1])
            torch.neg(xm[-2], out=df[-1])
        else:    # dft
            torch.sub(xm[1], xm[-1], out=df[0])
            torch.sub(xm[0], xm[-2], out=df[-1])
        out.append(df.div_(2 * v if v != 1 else 2.).movedim(0, d))
    return torch.stack(out, -1)


def grid_jacobian(grid, sample=None, bound="dft", voxel_size=1, type="grid", add_identity=True,
                  extrapolate=True):
    """Jacobian of a transformation / displacement field
    (nitorch/spatial/_grid.py:917-969): central finite differences, or the
    spline derivatives of grid_grad at `sample`.  Returns (..., *spatial,
    dim, dim) with jac[..., i, j] = d psi_i / d x_j."""
    grid = torch.as_tensor(grid)
    dim = grid.shape[-1]
    shape = grid.shape[-dim - 1:-1]
    if type == "grid":
        grid = grid - identity_grid(shape, dtype=grid.dtype, device=grid.device)
    if sample is None:
        jac = _diff_central(grid, list(range(-dim - 1, -1)), bound=bound, voxel_size=voxel_size)
    else:
        grid = grid.movedim(-1, -dim - 1)
        jac = grid_grad(grid, sample, bound=bound, extrapolate=extrapolate)
        jac = jac.movedim(-dim - 2, -2)
    if add_identity:
        torch.diagonal(jac, 0, -1, -2).add_(1)
    return jac
