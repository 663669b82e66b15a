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


class _Layout:
    """Where the caller's leading (batch) and channel dimensions go.

    The backend works on `(B, C, *spatial)` images and `(B, *spatial, D)`
    grids.  `lead` holds the broadcast leading dimensions that were folded
    into B, and `chan` the channel dimensions put back on the result: the
    input's own channel size when it had one, else a singleton channel when
    there were leading dimensions, else none."""

    __slots__ = ("lead", "chan", "dim")

    def __init__(self, lead, chan, dim):
        self.lead, self.chan, self.dim = tuple(lead), tuple(chan), dim

    def restore(self, out, vector=False):
        """(B, C, *spatial[, D]) backend result -> caller layout."""
        tail = tuple(out.shape[out.dim() - 1:]) if vector else ()
        stop = out.dim() - len(tail)
        return out.reshape(self.lead + self.chan + tuple(out.shape[stop - self.dim:stop]) + tail)


def _fold_grid(grid):
    """(..., *spatial, D) grid alone -> (B, *spatial, D) and its layout
    (nitorch/spatial/_grid.py:90-127, the count branch)."""
    dim = grid.shape[-1]
    lead = tuple(grid.shape[:grid.dim() - dim - 1])
    folded = grid.reshape((-1,) + tuple(grid.shape[grid.dim() - dim - 1:]))
    return folded, _Layout(lead, (1,) if lead else (), dim)


def _fold_pair(grid, image, shared_space=False):
    """Broadcast the leading dimensions of a grid (..., *gspatial, D) and an
    image (..., [C], *ispatial) and fold them into one batch dimension
    (nitorch/spatial/_grid.py:90-127).  With `shared_space` (push) the two
    spatial shapes are broadcast against each other as well.  Returns the
    (B, *spatial, D) grid, the (B, C, *spatial) image and the layout."""
    dim = grid.shape[-1]
    gspace = tuple(grid.shape[grid.dim() - dim - 1:-1])
    ispace = tuple(image.shape[image.dim() - dim:])
    has_chan = image.dim() > dim
    nchan = image.shape[image.dim() - dim - 1] if has_chan else 0
    ilead = tuple(image.shape[:image.dim() - dim - 1]) if has_chan else ()
    glead = tuple(grid.shape[:grid.dim() - dim - 1])
    if shared_space:
        gspace = ispace = tuple(torch.broadcast_shapes(gspace, ispace))
    lead = tuple(torch.broadcast_shapes(glead, ilead))
    width = nchan if nchan else 1
    grid = grid.expand(lead + gspace + (dim,)).reshape((-1,) + gspace + (dim,))
    image = image.expand(lead + (width,) + ispace).reshape((-1, width) + ispace)
    chan = (nchan,) if nchan else ((1,) if lead else ())
    return grid, image, _Layout(lead, chan, dim)


def _pull_labels(labels, grid, interpolation, bound, extrapolate, prefilter):
    """Label-map pull (nitorch/spatial/_grid.py:187-198): every label value
    is pulled as a soft indicator, and each output voxel keeps the label
    whose pulled indicator is strictly largest (the first label wins ties,
    voxels where every indicator is <= 0 keep 0).

    With `prefilter`, the reference assigns the in-place prefiltered
    indicator of one label to the image that the next labels' indicators
    are compared with (_grid.py:193-195); that statement is reproduced so
    that label maps resample as they do there."""
    dim = grid.shape[-1]
    shape = tuple(labels.shape[:2]) + tuple(grid.shape[1:-1])
    winner = labels.new_zeros(shape)
    best = grid.new_zeros(shape)
    source = labels
    for value in labels.unique():
        mask = (source == value).to(grid.dtype)
        if prefilter:
            source = spline_coeff_nd(mask, interpolation=interpolation, bound=bound, dim=dim,
                                     inplace=True)
        score = GridPull.apply(mask, grid, interpolation, bound, extrapolate)
        winner.masked_fill_(score > best, value)
        torch.maximum(best, score, out=best)
    return winner


def grid_pull(input, grid, interpolation="linear", bound="zero",
              extrapolate=False, prefilter=False):
    """Sample an image with respect to a deformation field
    (nitorch/spatial/_grid.py:146-205).

    input : (..., [channel], *inshape) tensor
    grid  : (..., *outshape, dim) tensor of voxel coordinates
    returns (..., [channel], *outshape)

    Integer inputs are label maps (see `_pull_labels`).  With `prefilter`,
    the input is first turned into interpolating spline coefficients
    (`spline_coeff_nd`)."""
    grid, image, layout = _fold_pair(grid, input)
    if not image.dtype.is_floating_point:
        return layout.restore(_pull_labels(image, grid, interpolation, bound, extrapolate, prefilter))
    if prefilter:
        image = spline_coeff_nd(image, interpolation=interpolation, bound=bound, dim=layout.dim)
    return layout.restore(GridPull.apply(image, grid, interpolation, bound, extrapolate))


def grid_push(input, grid, shape=None, interpolation="linear", bound="zero",
              extrapolate=False, prefilter=False):
    """Splat an image with respect to a deformation field (adjoint of pull,
    nitorch/spatial/_grid.py:208-251).  The output has `shape` (default: the
    broadcast spatial shape of input and grid); with `prefilter`, it is
    turned into interpolating spline coefficients in place."""
    grid, image, layout = _fold_pair(grid, input, shared_space=True)
    shape = tuple(image.shape[2:]) if shape is None else shape
    out = GridPush.apply(image, grid, shape, interpolation, bound, extrapolate)
    if prefilter:
        out = spline_coeff_nd(out, interpolation=interpolation, bound=bound, dim=layout.dim,
                              inplace=True)
    return layout.restore(out)


def grid_count(grid, shape=None, interpolation="linear", bound="zero",
               extrapolate=False):
    """Splatting weights, i.e. the push of an image of ones
    (nitorch/spatial/_grid.py:254-287); output shape `shape`, default the
    grid's spatial shape."""
    grid, layout = _fold_grid(grid)
    shape = tuple(grid.shape[1:-1]) if shape is None else shape
    return layout.restore(GridCount.apply(grid, shape, interpolation, bound, extrapolate))


def grid_grad(input, grid, interpolation="linear", bound="zero",
              extrapolate=False, prefilter=False):
    """Spatial gradients of the spline at the grid points
    (nitorch/spatial/_grid.py:290-328); returns (..., [channel], *outshape, dim).
    With `prefilter`, the input is first turned into interpolating spline
    coefficients."""
    grid, image, layout = _fold_pair(grid, input)
    if prefilter:
        image = spline_coeff_nd(image, interpolation=interpolation, bound=bound, dim=layout.dim)
    return layout.restore(GridGrad.apply(image, grid, interpolation, bound, extrapolate), vector=True)


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


def affine_grid(mat, shape, jitter=False):
    """Voxel coordinates `A @ [i, j, k] + t` of the lattice `shape` for an
    affine matrix [A | t] of shape (..., D[+1], D+1)
    (nitorch/spatial/_grid.py:579-615); returns (..., *shape, D), the
    matrix's leading dimensions first."""
    mat = torch.as_tensor(mat)
    ndim = mat.shape[-1] - 1
    if len(shape) != ndim:
        raise ValueError(f"affine_grid: a ({ndim}+1)-column matrix needs a {ndim}-D shape, got {list(shape)}")
    if mat.shape[-2] not in (ndim, ndim + 1):
        raise ValueError(f"affine_grid: expected (..., {ndim}[+1], {ndim + 1}) matrices, got {tuple(mat.shape)}")
    lead = tuple(mat.shape[:-2])
    ijk = identity_grid(list(shape), mat.dtype, mat.device, jitter=jitter)
    ijk = ijk.reshape((1,) * len(lead) + tuple(ijk.shape))
    lin = mat[..., :ndim, :ndim].reshape(lead + (1,) * ndim + (ndim, ndim))
    off = mat[..., :ndim, ndim].reshape(lead + (1,) * ndim + (ndim,))
    return _matvec(lin, ijk) + off


def _resize_lattice(image, factor, shape, anchor, dim):
    """The output shape and the per-axis `LatticeAxis` maps of a resize.

    The output length of an axis is `shape` when given, else
    max(1, int(n * factor)); a factor left unset is output / input length
    (nitorch/spatial/_grid.py:693-700)."""
    from .affine import LatticeAxis, _per_axis
    if dim is None:
        dim = len(shape) if shape is not None else image.dim() - 2
    n_in = list(image.shape[image.dim() - dim:])
    if torch.is_tensor(factor):
        factor = factor.tolist()
    factor = [float(f or 0.) for f in _per_axis(factor if factor is not None else 0., dim)]
    wanted = _per_axis(shape if shape is not None else [None], dim)
    n_out = [m or max(1, int(n * f)) for n, m, f in zip(n_in, wanted, factor)]
    if not all(n_out):
        raise ValueError("Either factor or shape must be set in all dimensions")
    factor = [f or m / n for n, m, f in zip(n_in, n_out, factor)]
    anchor = _per_axis(anchor, dim)
    axes = [LatticeAxis(a, f, n, m) for a, f, n, m in zip(anchor, factor, n_in, n_out)]
    return n_in, n_out, factor, anchor, axes


def resize(image, factor=None, shape=None, affine=None, anchor="c", dim=None, *args, **kwargs):
    """Resize an image by a factor or to a shape (nitorch/spatial/_grid.py:618-754).

    The image is sampled by `grid_pull` (prefilter=True unless given; other
    arguments go to it) on the lattice whose axes `LatticeAxis` describes:
    anchor 'c' (voxel centres), 'e' (outer edges), 'f' (first voxel) or 'l'
    (last voxel), per dimension.  Returns the resized image and, when
    `affine` is given, the resized lattice's orientation matrix."""
    from .affine import affine_resize
    image = torch.as_tensor(image)
    with_map = kwargs.pop("_return_trf", False)
    n_in, n_out, factor, anchor, axes = _resize_lattice(image, factor, shape, anchor, dim)
    lines = [ax.coordinates(dtype=image.dtype, device=image.device) for ax in axes]
    lattice = torch.stack(torch.meshgrid(*lines, indexing="ij"), dim=-1)
    kwargs.setdefault("prefilter", True)
    result = [grid_pull(image, lattice, *args, **kwargs)]
    if affine is not None:
        result.append(affine_resize(affine, n_in, factor, anchor, shape_out=n_out)[0])
    if with_map:
        result.append(([ax.scale for ax in axes], [ax.shift for ax in axes]))
    return result[0] if len(result) == 1 else tuple(result)


def resize_grid(grid, factor=None, shape=None, type="grid", affine=None, *args, **kwargs):
    """Resize a transformation (type 'grid') or displacement field and
    express it in the resized voxel space (nitorch/spatial/_grid.py:757-858).

    Each component is resampled like an image, then mapped from input to
    output voxel units by the inverse of the axis map `scale * o + shift`:
    (v - shift) / scale for coordinates, v / scale for displacements."""
    kwargs["_return_trf"] = True
    kwargs.setdefault("prefilter", True)
    ndim = grid.shape[-1]
    lead = tuple(grid.shape[:grid.dim() - ndim - 1])
    fields = grid.reshape((-1,) + tuple(grid.shape[grid.dim() - ndim - 1:])).movedim(-1, 1)
    if affine is not None:
        affine = affine.expand(lead + tuple(affine.shape[-2:])).reshape((-1,) + tuple(affine.shape[-2:]))
        fields, affine, (scales, shifts) = resize(fields, factor, shape, affine, *args, **kwargs)
    else:
        fields, (scales, shifts) = resize(fields, factor, shape, affine, *args, **kwargs)
    is_coord = type[0].lower() == "g"
    parts = [(fields[:, d] - shifts[d] if is_coord else fields[:, d]) / scales[d] for d in range(ndim)]
    out = torch.stack(parts, -1)
    out = out.reshape(lead + tuple(out.shape[1:]))
    if affine is not None:
        return out, affine.reshape(lead + tuple(affine.shape[-2:]))
    return out


def reslice(image, affine, affine_to, shape_to=None, **kwargs):
    """Resample an image into another space (shape + affine)
    (nitorch/spatial/_grid.py:861-915): the target voxel v is read at the
    source voxel inv(affine) @ affine_to @ v, by `grid_pull` with its
    options (prefilter=True unless given).  Image (*batch, *channels,
    *spatial) with batch the affine's leading dimensions; returns
    (*batch, *channels, *shape_to)."""
    from .affine import affine_lmdiv
    kwargs.setdefault("prefilter", True)
    image = torch.as_tensor(image)
    dtype = image.dtype if image.dtype.is_floating_point else torch.get_default_dtype()
    affine = torch.as_tensor(affine).to(dtype=dtype, device=image.device)
    affine_to = torch.as_tensor(affine_to).to(dtype=dtype, device=image.device)
    dim = affine.shape[-1] - 1
    lead = tuple(affine.shape[:-2])
    chans = tuple(image.shape[len(lead):image.dim() - dim])
    space = tuple(image.shape[image.dim() - dim:])
    shape_to = tuple(space if shape_to is None else shape_to)
    nbatch = int(torch.Size(lead).numel()) if lead else 0
    nchan = int(torch.Size(chans).numel()) if chans else 0
    lattice = affine_grid(affine_lmdiv(affine, affine_to), shape_to)
    if not nbatch:
        lattice = lattice[None]
    out = grid_pull(image.reshape((nbatch or 1, nchan or 1) + space), lattice, **kwargs)
    return out.reshape(lead + chans + shape_to)


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
