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


def affine_resize(affine, shape, factor, anchor="c", shape_out=None):
    """Affine of a resized lattice (nitorch/spatial/_affine.py:1712-1790):
    returns (affine, shape_out)."""
    affine = torch.as_tensor(affine)
    nb_dim = affine.shape[-1] - 1
    factor = list(factor) if isinstance(factor, (list, tuple)) else [factor]
    factor = [float(f) for f in (factor + factor[-1:] * nb_dim)[:nb_dim]]
    anchor = list(anchor) if isinstance(anchor, (list, tuple)) else [anchor]
    anchor = [a[0].lower() for a in (anchor + anchor[-1:] * nb_dim)[:nb_dim]]
    if shape_out is None:
        shape_out = [max(1, int(round(s * f))) for s, f in zip(shape, factor)]
    shifts, scales = [], []
    for anch, f, inshp, outshp in zip(anchor, factor, shape, shape_out):
        if inshp == 1 or outshp == 1:
            anch = "e"
        if anch == "c":
            shifts.append(0.)
            scales.append((inshp - 1) / (outshp - 1))
        elif anch == "e":
            shifts.append(0.5 * (inshp / outshp - 1))
            scales.append(inshp / outshp)
        elif anch == "f":
            shifts.append(0.)
            scales.append(1 / f)
        elif anch == "l":
            shifts.append((inshp - 1) - (outshp - 1) / f)
            scales.append(1 / f)
        else:
            raise ValueError("Unknown anchor {}".format(anch))
        if inshp == outshp == 1:
            scales[-1] = 1 / f
    lin = torch.diag(torch.as_tensor(scales, dtype=affine.dtype, device=affine.device))
    trl = torch.as_tensor(shifts, dtype=affine.dtype, device=affine.device)[..., None]
    trf = torch.cat((lin, trl), dim=1)
    return affine_matmul(affine, trf), tuple(shape_out)
