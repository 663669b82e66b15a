"""Affine-matrix algebra used by `resize` / `reslice`
(nitorch/spatial/_affine.py: `affine_matvec` :1568-1606, `affine_matmul`
:1608-1639, `affine_inv` :1641-1670, `affine_lmdiv` :1672-1690,
`affine_resize` :1712-1790).

Matrices are (..., D[+1], D+1): square ones are homogeneous (last row
[0, ..., 0, 1]) and results keep the homogeneity of the left operand, as
the reference's `is_homogeneous` attribute does by default.  These are
small host-side torch operations on the matrices, not on images.
"""
import torch


def _euclidean(a):
    a = torch.as_tensor(a)
    return a[..., :-1, :] if a.shape[-1] == a.shape[-2] else a


def _homogeneous(a):
    last = torch.zeros(a.shape[:-2] + (1, a.shape[-1]), dtype=a.dtype, device=a.device)
    last[..., -1] = 1
    return torch.cat((a, last), dim=-2)


def affine_matvec(affine, vector):
    """affine @ vector for (..., D[+1]) vectors (homogeneous kept as such)."""
    affine = _euclidean(affine)
    vector = torch.as_tensor(vector)
    affine = affine.to(dtype=vector.dtype, device=vector.device)
    nd = affine.shape[-1] - 1
    is_h = vector.shape[-1] == nd + 1
    out = torch.matmul(affine[..., :, :-1], vector[..., :nd, None])[..., 0] + affine[..., :, -1]
    if is_h:
        out = torch.cat((out, torch.ones(out.shape[:-1] + (1,), dtype=out.dtype, device=out.device)), -1)
    return out


def affine_matmul(a, b):
    """a @ b of affine matrices; homogeneous iff `a` is square."""
    a = torch.as_tensor(a)
    b = torch.as_tensor(b)
    is_h = a.shape[-1] == a.shape[-2]
    a, b = _euclidean(a), _euclidean(b)
    dtype = torch.promote_types(a.dtype, b.dtype)
    a, b = a.to(dtype), b.to(dtype=dtype, device=a.device)
    Z = torch.matmul(a[..., :, :-1], b[..., :, :-1])
    T = torch.matmul(a[..., :, :-1], b[..., :, -1:])[..., 0] + a[..., :, -1]
    out = torch.cat((Z, T[..., None]), dim=-1)
    return _homogeneous(out) if is_h else out


def affine_inv(affine):
    """Inverse (pseudo-inverse if rectangular) of an affine matrix."""
    affine = torch.as_tensor(affine)
    is_h = affine.shape[-1] == affine.shape[-2]
    a = _euclidean(affine)
    nin, nout = a.shape[-1] - 1, a.shape[-2]
    inverse = torch.inverse if nin == nout else torch.pinverse
    zoom = inverse(a[..., :, :-1])
    trl = -torch.matmul(zoom, a[..., :, -1:])[..., 0]
    out = torch.cat((zoom, trl[..., None]), dim=-1)
    return _homogeneous(out) if is_h else out


def affine_lmdiv(a, b):
    """inv(a) @ b."""
    return affine_matmul(affine_inv(a), b)


class LatticeAxis:
    """One axis of a resized lattice: output voxel o sits at input voxel
    coordinate `scale * o + shift`.

    The anchor decides which points of the two lattices coincide
    (nitorch/spatial/_affine.py:1712-1790, nitorch/spatial/_grid.py:618-754):
    'c' the centres of the first and last voxels, 'e' the outer edges of the
    volume, 'f' the centre of the first voxel and 'l' that of the last one,
    the last two with the voxel size divided exactly by `factor`.  An axis
    of length 1 on either side is always edge-anchored, and when both sides
    have length 1 the scale is 1 / factor.
    """

    __slots__ = ("anchor", "factor", "n_in", "n_out", "scale", "shift")

    def __init__(self, anchor, factor, n_in, n_out):
        anchor = anchor[0].lower()
        if anchor not in "cefl":
            raise ValueError(f"Unknown anchor {anchor}")
        if n_in == 1 or n_out == 1:
            anchor = "e"
        self.anchor, self.factor, self.n_in, self.n_out = anchor, factor, n_in, n_out
        if anchor == "c":
            self.scale, self.shift = (n_in - 1) / (n_out - 1), 0.
        elif anchor == "e":
            self.scale = n_in / n_out
            self.shift = 0.5 * (self.scale - 1)
        else:
            self.scale = 1 / factor
            self.shift = 0. if anchor == "f" else (n_in - 1) - (n_out - 1) / factor
        if n_in == n_out == 1:
            self.scale = 1 / factor

    def coordinates(self, dtype=None, device=None):
        """Input coordinates of the output voxels, as a 1-D tensor.  The
        arithmetic follows the reference's so that sampled values agree to
        the last bit: `linspace` for centres, a product with the scale for
        edges, a division by the factor for first / last."""
        if self.anchor == "c":
            return torch.linspace(0, self.n_in - 1, self.n_out, dtype=dtype, device=device)
        o = torch.arange(0., self.n_out, dtype=dtype, device=device)
        if self.anchor == "e":
            return o * self.scale + self.shift
        o = o / self.factor
        return o + self.shift if self.anchor == "l" else o


def _per_axis(values, n):
    """A value or a sequence, extended to `n` entries by repeating its last."""
    values = list(values) if isinstance(values, (list, tuple)) else [values]
    return values + values[-1:] * (n - len(values)) if len(values) < n else values[:n]


def affine_resize(affine, shape, factor, anchor="c", shape_out=None):
    """Orientation matrix of a resized lattice, and the resized shape
    (nitorch/spatial/_affine.py:1712-1790).  The new voxel-to-world matrix
    is `affine @ [diag(scale) | shift]`, with each axis's scale and shift
    from `LatticeAxis`; `shape_out` defaults to round(shape * factor), at
    least 1."""
    affine = torch.as_tensor(affine)
    ndim = affine.shape[-1] - 1
    factor = [float(f) for f in _per_axis(factor, ndim)]
    anchor = _per_axis(anchor, ndim)
    if shape_out is None:
        shape_out = [max(1, int(round(n * f))) for n, f in zip(shape, factor)]
    axes = [LatticeAxis(a, f, n, m) for a, f, n, m in zip(anchor, factor, shape, shape_out)]
    voxel_map = torch.zeros((ndim, ndim + 1), dtype=affine.dtype, device=affine.device)
    for d, ax in enumerate(axes):
        voxel_map[d, d] = ax.scale
        voxel_map[d, ndim] = ax.shift
    return affine_matmul(affine, voxel_map), tuple(shape_out)
