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
Hessians stored sparse as (h00, h11, h22, h01, h02, h12).  3-D fields of
3 x 3 matrices on the GPU are multiplied, and pushed forward through a
squaring, by libnitorch_amd.so (`ni_svf_mat3_mul`, `ni_svf_grad_step`),
reading transposed and channel-first views in place.
"""
import torch

from . import _lib as L
from . import spatial as SP

L.register_signatures({
    "ni_svf_mat3_mul": [L._I32, L._I64, L._P, L._PI64, L._P, L._PI64, L._P, L._P],
    "ni_svf_grad_step": [L._I32, L._I64, L._P, L._PI64, L._P, L._PI64, L._P, L._PI64, L._P, L._P],
})


def _vox_stride(t, trailing):
    """Element stride of the voxel index when all but the `trailing` dims of
    t collapse into one (None if they do not)."""
    lead = [(n, st) for n, st in zip(t.shape[:t.dim() - trailing], t.stride()[:t.dim() - trailing]) if n != 1]
    if not lead:
        return 0
    step = lead[-1][1]
    span = step * lead[-1][0]
    for n, st in reversed(lead[:-1]):
        if st != span:
            return None
        span *= n
    return step


def _native3(*ts):
    """The GPU kernels apply: 3-D float32 / float64 CUDA fields of one shape."""
    t0 = ts[0]
    return (t0.is_cuda and t0.dtype in (torch.float32, torch.float64) and t0.shape[-1] == 3
            and all(t.device == t0.device and t.dtype == t0.dtype for t in ts))


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


# Products of fields of small matrices as sums of d rank-one terms, each
# one elementwise kernel over the output (in-place multiply-adds), rather
# than batched GEMMs of millions of 3 x 3s or a d-times wider temporary
# reduced over a strided axis.
def _matvec(mat, vec):
    out = mat[..., :, 0] * vec[..., 0:1]
    for k in range(1, mat.shape[-1]):
        out.addcmul_(mat[..., :, k], vec[..., k:k + 1])
    return out


def _mat3_into(a, b, out):
    """out = a @ b on the GPU; leading dims that do not collapse into one
    voxel stride are split.  False if a view cannot be addressed."""
    sa, sb = _vox_stride(a, 2), _vox_stride(b, 2)
    if sa is None or sb is None:
        return a.dim() > 3 and all(_mat3_into(a[i], b[i], out[i]) for i in range(a.shape[0]))
    L.check(L.lib().ni_svf_mat3_mul(L.dtype_code(a.dtype), out.numel() // 9, L.ptr(a),
                                    L.i64arr([sa, a.stride(-2), a.stride(-1)]), L.ptr(b),
                                    L.i64arr([sb, b.stride(-2), b.stride(-1)]), L.ptr(out),
                                    L.stream_of(a)), "svf_mat3_mul")
    return True


def _matmul(a, b):
    if _native3(a, b) and a.shape == b.shape and a.shape[-2] == 3:
        out = torch.empty(a.shape, dtype=a.dtype, device=a.device)
        if _mat3_into(a, b, out):
            return out
    out = a[..., :, 0:1] * b[..., 0:1, :]
    for k in range(1, a.shape[-1]):
        out.addcmul_(a[..., :, k:k + 1], b[..., k:k + 1, :])
    return out


def _step_into(jac, p, g0, out):
    sj, sp, sg = _vox_stride(jac, 2), _vox_stride(p, 1), _vox_stride(g0, 1)
    if sj is None or sp is None or sg is None:
        return p.dim() > 2 and all(_step_into(jac[i], p[i], g0[i], out[i]) for i in range(p.shape[0]))
    L.check(L.lib().ni_svf_grad_step(L.dtype_code(jac.dtype), out.numel() // 3, L.ptr(jac),
                                     L.i64arr([sj, jac.stride(-2), jac.stride(-1)]), L.ptr(p),
                                     L.i64arr([sp, p.stride(-1)]), L.ptr(g0), L.i64arr([sg, g0.stride(-1)]),
                                     L.ptr(out), L.stream_of(g0)), "svf_grad_step")
    return True


def _grad_step(jac, p, g0):
    """g0 + det(jac) jac^T p (one squaring's push-forward of the gradient,
    _svf.py:272-279), fused for 3-D GPU fields; None if not applicable."""
    if not (_native3(jac, p, g0) and jac.shape[:-1] == p.shape == g0.shape and jac.shape[-2] == 3):
        return None
    out = torch.empty(g0.shape, dtype=g0.dtype, device=g0.device)
    return out if _step_into(jac, p, g0, out) else None


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
        grid = ident + vel
        if jacobian:
            jac = _composition_jac(jac, grid, lhs=vel, type=("grid", "displacement"))
        vel = iadd(vel, _pull_vel(vel, grid, interpolation=interpolation, bound=bound))
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
        grid = ident + vel    # the squaring's transformation, shared by its pulls
        pulled = _pull_vel(grad, grid, **opt)
        step = _grad_step(jac, pulled, grad) if hess is None else None
        if step is not None:
            grad = step
        else:
            det = _det(jac)
            jac = jac.transpose(-1, -2)
            grad0 = grad
            grad = _matvec(jac, pulled)
            grad *= det[..., None]
            grad += grad0
            if hess is not None:
                hess0 = hess
                hess = _pull_vel(hess, grid, **opt)
                hess = _jhj(jac, hess)
                hess *= det[..., None]
                hess += hess0
            jac = jac.transpose(-1, -2)
        jac = _composition_jac(jac, grid, lhs=vel, type=("grid", "disp"), **opt)
        vel += _pull_vel(vel, grid, **opt)
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
