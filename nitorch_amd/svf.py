"""Exponentiation of stationary velocity fields by scaling and squaring.

Restates nitorch/spatial/_svf.py (`exp` :11-42, `exp_forward` :45-99,
`_jhj` :102-186, `exp_backward` :189-297, `_Exp` :300-320, `_pull_vel`
:323-361, `_push_vel` :364-402, `_pull_jac` :405-446, `_composition_jac`
:449-478) on this package's grid_pull / grid_push: the velocity is divided
by 2**steps and composed with itself `steps` times, v <- v + v o (id + v),
each composition one order-`order` pull of the 3-channel field
(extrapolate, dft by default).  Gradients with respect to the velocity flow
either through the pull's backward kernels (autodiff, `anagrad=False`) or
through `exp_backward`, the DARTEL-style push-forward of the gradient across
the squarings (`anagrad=True`).

Fields are channel-last like the reference: ([batch], *spatial, dim);
Jacobians (..., dim, dim) with jac[..., i, j] = d psi_i / d x_j; symmetric
Hessians stored sparse as (h00, h11, h22, h01, h02, h12).
"""
import torch

from . import spatial as SP


def _pull_vel(vel, grid, interpolation="linear", bound="dft", extrapolate=True):
    """Interpolate a ([batch], *spatial, dim) field at a ([batch], *spatial,
    dim) transformation (nitorch/spatial/_svf.py:323-361)."""
    dim = grid.shape[-1]
    vel = vel.movedim(-1, -dim - 1)
    vel_nb, grid_nb = vel.dim() == dim + 1, grid.dim() == dim + 1
    if vel_nb:
        vel = vel[None]
    if grid_nb:
        grid = grid[None]
    out = SP.grid_pull(vel, grid, interpolation=interpolation, bound=bound, extrapolate=extrapolate)
    out = out.movedim(-dim - 1, -1)
    return out[0] if (vel_nb and grid_nb) else out


def _push_vel(vel, grid, interpolation="linear", bound="dft", extrapolate=True):
    """Splat a ([batch], *spatial, dim) field through a transformation
    (nitorch/spatial/_svf.py:364-402)."""
    dim = grid.shape[-1]
    vel = vel.movedim(-1, -dim - 1)
    vel_nb, grid_nb = vel.dim() == dim + 1, grid.dim() == dim + 1
    if vel_nb:
        vel = vel[None]
    if grid_nb:
        grid = grid[None]
    out = SP.grid_push(vel, grid, interpolation=interpolation, bound=bound, extrapolate=extrapolate)
    out = out.movedim(-dim - 1, -1)
    return out[0] if (vel_nb and grid_nb) else out


def _pull_jac(jac, grid, **kwargs):
    """Pull each entry of a ([batch], *spatial, dim, dim) Jacobian field
    (linear, dft, extrapolated by default; _svf.py:405-446)."""
    kwargs.setdefault("bound", "dft")
    kwargs.setdefault("extrapolate", True)
    dim = grid.shape[-1]
    j = jac.reshape([*jac.shape[:-2], -1]).movedim(-1, -dim - 1)
    j_nb, g_nb = j.dim() == dim + 1, grid.dim() == dim + 1
    if j_nb:
        j = j[None]
    if g_nb:
        grid = grid[None]
    j = SP.grid_pull(j, grid, **kwargs).movedim(-dim - 1, -1)
    j = j.reshape([*j.shape[:-1], dim, dim])
    return j[0] if (j_nb and g_nb) else j


def _composition_jac(jac, rhs, lhs=None, type="grid", identity=None, **kwargs):
    """Jacobian of lhs o rhs given the Jacobian `jac` of rhs (_svf.py:449-478):
    the left factor's central-difference Jacobian (dft) pulled (linear, dft)
    to rhs, times jac.  Extra keywords are accepted and unused, as there."""
    if lhs is None:
        lhs = rhs
    dim = rhs.shape[-1]
    typer, typel = (list(type) + [type[-1]])[:2] if isinstance(type, (list, tuple)) else (type, type)
    jac_left = SP.grid_jacobian(lhs, type=typel)
    if typer != "grid":
        if identity is None:
            identity = SP.identity_grid(rhs.shape[-dim - 1:-1], dtype=rhs.dtype, device=rhs.device)
        rhs = rhs + identity
    jac_left = _pull_jac(jac_left, rhs)
    return _matmul(jac_left, jac)


# Products of fields of small matrices as broadcast multiply-and-sum (one
# elementwise kernel each) rather than batched GEMMs of millions of 3 x 3s.
def _matvec(mat, vec):
    return (mat * vec[..., None, :]).sum(-1)


def _matmul(a, b):
    return (a[..., :, :, None] * b[..., None, :, :]).sum(-2)


def _det(m):
    """Determinant of (..., d, d) matrices; closed form for d = 3."""
    if m.shape[-1] != 3:
        return torch.linalg.det(m)
    a, b, c = m[..., 0, 0], m[..., 0, 1], m[..., 0, 2]
    d, e, f = m[..., 1, 0], m[..., 1, 1], m[..., 1, 2]
    g, h, i = m[..., 2, 0], m[..., 2, 1], m[..., 2, 2]
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)


def _inv(m):
    """Inverse of (..., d, d) matrices; adjugate / determinant for d = 3."""
    if m.shape[-1] != 3:
        return torch.linalg.inv(m)
    a, b, c = m[..., 0, 0], m[..., 0, 1], m[..., 0, 2]
    d, e, f = m[..., 1, 0], m[..., 1, 1], m[..., 1, 2]
    g, h, i = m[..., 2, 0], m[..., 2, 1], m[..., 2, 2]
    adj = torch.stack([torch.stack([e * i - f * h, c * h - b * i, b * f - c * e], -1),
                       torch.stack([f * g - d * i, a * i - c * g, c * d - a * f], -1),
                       torch.stack([d * h - e * g, b * g - a * h, a * e - b * d], -1)], -2)
    return adj / _det(m)[..., None, None]


def _jhj(jac, hess):
    """J H J^T of a sparse symmetric Hessian (..., dim*(dim+1)/2), same
    storage order out (_svf.py:102-186)."""
    dim = jac.shape[-1]
    if dim == 1:
        return jac[..., 0, 0:1] * jac[..., 0, 0:1] * hess
    if dim == 2:
        h00, h11, h01 = hess[..., 0], hess[..., 1], hess[..., 2]
        j00, j01, j10, j11 = jac[..., 0, 0], jac[..., 0, 1], jac[..., 1, 0], jac[..., 1, 1]
        return torch.stack([j00 * j00 * h00 + j01 * j01 * h11 + 2 * j00 * j01 * h01,
                            j10 * j10 * h00 + j11 * j11 * h11 + 2 * j10 * j11 * h01,
                            j00 * j10 * h00 + j01 * j11 * h11 + (j01 * j10 + j00 * j11) * h01], -1)
    H = torch.stack([torch.stack([hess[..., 0], hess[..., 3], hess[..., 4]], -1),
                     torch.stack([hess[..., 3], hess[..., 1], hess[..., 5]], -1),
                     torch.stack([hess[..., 4], hess[..., 5], hess[..., 2]], -1)], -2)
    out = _matmul(_matmul(jac, H), jac.transpose(-1, -2))
    return torch.stack([out[..., 0, 0], out[..., 1, 1], out[..., 2, 2],
                        out[..., 0, 1], out[..., 0, 2], out[..., 1, 2]], -1)


def exp_forward(vel, inverse=False, steps=8, interpolation="linear", bound="dft", displacement=False,
                jacobian=False, _anagrad=False):
    """Scaling and squaring (nitorch/spatial/_svf.py:45-99): returns the
    transformation (or, with `displacement`, the displacement) exp(vel), and
    with `jacobian` its Jacobian accumulated over the squarings."""
    vel = -vel if inverse else vel.clone()
    dim = vel.shape[-1]
    spatial = vel.shape[-1 - dim:-1]
    ident = SP.identity_grid(spatial, dtype=vel.dtype, device=vel.device)
    jac = torch.eye(dim, dtype=vel.dtype, device=vel.device).expand([*vel.shape[:-1], dim, dim])
    iadd = (lambda x, y: x.add(y)) if (not _anagrad and vel.requires_grad) else (lambda x, y: x.add_(y))
    vel = vel / (2 ** steps)
    for _ in range(steps):
        if jacobian:
            jac = _composition_jac(jac, vel, type="displacement", identity=ident)
        vel = iadd(vel, _pull_vel(vel, ident + vel, interpolation=interpolation, bound=bound))
    if not displacement:
        vel = vel + ident
    return (vel, jac) if jacobian else vel


def exp_backward(vel, *gradhess, inverse=False, steps=8, interpolation="linear", bound="dft",
                 rotate_grad=False):
    """Gradient (and block-diagonal Hessian) with respect to the velocity
    from those with respect to exp(vel) (nitorch/spatial/_svf.py:189-297):
    the gradient is pushed forward through each squaring,
    g <- g + |J| J^T g(id + v), scaled by 2^-steps; `rotate_grad` first
    rotates it by the Jacobian of the forward exponentiation."""
    grad, *rest = gradhess
    hess = rest[0] if rest else None
    opt = dict(bound=bound, interpolation=interpolation)
    dim = vel.shape[-1]
    shape = vel.shape[-dim - 1:-1]
    ident = SP.identity_grid(shape, dtype=vel.dtype, device=vel.device)
    vel = vel.detach().clone()
    grad = grad.detach()
    hess = hess.detach() if hess is not None else None
    if rotate_grad:
        _, jac = exp_forward(vel, jacobian=True, steps=steps, displacement=True, _anagrad=True, **opt)
        jac = jac.transpose(-1, -2)
        grad = _matvec(jac, grad)
        if hess is not None:
            hess = _jhj(jac, hess)
        del jac
    vel /= (-1 if not inverse else 1) * (2 ** steps)
    jac = SP.grid_jacobian(vel, bound=bound, type="disp")
    # rotate the gradient so that steps == 0 gives the small-deformation one
    ijac = 2 * torch.eye(dim, dtype=jac.dtype, device=jac.device) - jac
    ijac = _inv(ijac.transpose(-1, -2))
    grad = _matvec(ijac, grad)
    del ijac
    for _ in range(steps):
        det = _det(jac)
        jac = jac.transpose(-1, -2)
        grad0 = grad
        grad = _pull_vel(grad, ident + vel, **opt)
        grad = _matvec(jac, grad)
        grad *= det[..., None]
        grad += grad0
        if hess is not None:
            hess0 = hess
            hess = _pull_vel(hess, ident + vel, **opt)
            hess = _jhj(jac, hess)
            hess *= det[..., None]
            hess += hess0
        jac = jac.transpose(-1, -2)
        jac = _composition_jac(jac, vel, type="disp", identity=ident, **opt)
        vel += _pull_vel(vel, ident + vel, **opt)
    if inverse:
        grad.neg_()
    grad /= (2 ** steps)
    if hess is not None:
        hess /= (2 ** steps)
    return (grad, hess) if hess is not None else grad


class _Exp(torch.autograd.Function):
    """exp with the analytic backward (nitorch/spatial/_svf.py:300-320)."""

    @staticmethod
    def forward(ctx, vel, inverse, steps, interpolation, bound, displacement):
        if vel.requires_grad:
            ctx.save_for_backward(vel)
            ctx.args = dict(steps=steps, inverse=inverse, interpolation=interpolation, bound=bound)
        return exp_forward(vel, inverse, steps, interpolation, bound, displacement, _anagrad=True)

    @staticmethod
    def backward(ctx, grad):
        vel, = ctx.saved_tensors
        grad = exp_backward(vel, grad, rotate_grad=True, **ctx.args)
        return (grad,) + (None,) * 5


def exp(vel, inverse=False, steps=8, interpolation="linear", bound="dft", displacement=False,
        anagrad=False):
    """Exponentiate a stationary velocity field (nitorch/spatial/_svf.py:11-42):
    autodiff gradients through the HIP pull kernels, or with `anagrad` the
    analytic backward `exp_backward`."""
    exp_fn = _Exp.apply if anagrad else exp_forward
    return exp_fn(vel, inverse, steps, interpolation, bound, displacement)
