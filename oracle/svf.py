"""ORACLE — test infrastructure only.

FP64 numpy restatement of scaling and squaring of a stationary velocity
field (nitorch/spatial/_svf.py: exp_forward :45-99 with its Jacobian,
_jhj :102-186, exp_backward :189-297, _pull_vel :323-361, _pull_jac
:405-446, _composition_jac :449-478) and of the central finite differences
behind grid_jacobian (nitorch/spatial/_grid.py:917-969,
_finite_differences.py:32-221), on the oracle's pull (`pushpull_oracle.c`,
extrapolated).  Fields are channel-last: vel [N, *spatial, 3]; Jacobians
[N, *spatial, 3, 3] with jac[..., i, j] = d psi_i / d x_j; sparse symmetric
Hessians [N, *spatial, 6] ordered h00 h11 h22 h01 h02 h12.
"""
import numpy as np

from . import pushpull as PP

_ORDERS = {"nearest": 0, "linear": 1, "quadratic": 2, "cubic": 3}


def identity(shape):
    return np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij"), -1)


def pull_vel(vel, grid, order=1, bound="dft"):
    """vel, grid [N, *spatial, 3] -> vel(grid) [N, *spatial, 3]."""
    v = np.ascontiguousarray(np.moveaxis(vel, -1, 1))
    out = PP.grid_pull(v, np.ascontiguousarray(grid), bound, order, 1)
    return np.moveaxis(out, 1, -1)


def exp_forward(vel, inverse=False, steps=8, interpolation="linear", bound="dft", displacement=False):
    order = _ORDERS.get(interpolation, interpolation)
    vel = np.asarray(vel, np.float64)
    vel = -vel if inverse else vel.copy()
    ident = identity(vel.shape[1:-1])[None]
    vel = vel / (2 ** steps)
    for _ in range(steps):
        vel = vel + pull_vel(vel, ident + vel, order, bound)
    return vel if displacement else vel + ident


def diff_central(x, axes, bound):
    """First-order central differences along `axes`, stacked last
    (_finite_differences.py:154-184): (x[i+1] - x[i-1]) / 2 inside, the
    bound's rule on the two end samples, and a copy of x along an axis of
    length 1 (as the reference returns)."""
    out = []
    for ax in axes:
        xm = np.moveaxis(x, ax, 0)
        n = xm.shape[0]
        if n == 1:
            out.append(np.moveaxis(xm.copy(), 0, ax))
            continue
        d = np.empty_like(xm)
        d[1:-1] = xm[2:] - xm[:-2]
        if bound in ("dct2", "replicate"):
            d[0] = xm[1] - xm[0]
            d[-1] = xm[-1] - xm[-2]
        elif bound == "dct1":
            d[0] = 0
            d[-1] = 0
        elif bound == "dst2":
            d[0] = xm[1] + xm[0]
            d[-1] = -(xm[-1] + xm[-2])
        elif bound in ("dst1", "zero"):
            d[0] = xm[1]
            d[-1] = -xm[-2]
        else:
            d[0] = xm[1] - xm[-1]
            d[-1] = xm[0] - xm[-2]
        out.append(np.moveaxis(d / 2, 0, ax))
    return np.stack(out, -1)


def grid_jacobian(grid, bound="dft", type="grid", add_identity=True):
    """[N, *spatial, 3] -> [N, *spatial, 3, 3] by central differences."""
    grid = np.asarray(grid, np.float64)
    dim = grid.shape[-1]
    if type == "grid":
        grid = grid - identity(grid.shape[-dim - 1:-1])
    jac = diff_central(grid, list(range(grid.ndim - dim - 1, grid.ndim - 1)), bound)
    if add_identity:
        jac = jac + np.eye(dim)
    return jac


def pull_jac(jac, grid):
    """Linear, dft, extrapolated pull of each Jacobian entry (_svf.py:405-446;
    _composition_jac passes no options)."""
    shp = jac.shape
    j = jac.reshape(shp[:-2] + (-1,))
    j = np.ascontiguousarray(np.moveaxis(j, -1, 1))
    out = PP.grid_pull(j, np.ascontiguousarray(grid), "dft", 1, 1)
    return np.moveaxis(out, 1, -1).reshape(grid.shape[:-1] + shp[-2:])


def composition_jac(jac, rhs, displacement=True):
    """Jacobian of rhs o rhs given jac of the right factor (_svf.py:449-478)."""
    left = grid_jacobian(rhs, bound="dft", type="disp" if displacement else "grid")
    if displacement:
        rhs = rhs + identity(rhs.shape[1:-1])[None]
    left = pull_jac(left, rhs)
    return left @ jac


def _sym(h):
    H = np.empty(h.shape[:-1] + (3, 3))
    H[..., 0, 0], H[..., 1, 1], H[..., 2, 2] = h[..., 0], h[..., 1], h[..., 2]
    H[..., 0, 1] = H[..., 1, 0] = h[..., 3]
    H[..., 0, 2] = H[..., 2, 0] = h[..., 4]
    H[..., 1, 2] = H[..., 2, 1] = h[..., 5]
    return H


def _unsym(H):
    return np.stack([H[..., 0, 0], H[..., 1, 1], H[..., 2, 2], H[..., 0, 1], H[..., 0, 2], H[..., 1, 2]], -1)


def jhj(jac, hess):
    """J H J^T of a sparse symmetric Hessian (_svf.py:102-186), densely."""
    return _unsym(jac @ _sym(hess) @ np.swapaxes(jac, -1, -2))


def exp_forward_jac(vel, steps=8, interpolation="linear", bound="dft"):
    """exp_forward(vel, jacobian=True, displacement=True): (disp, jac)."""
    order = _ORDERS.get(interpolation, interpolation)
    vel = np.asarray(vel, np.float64) / (2 ** steps)
    ident = identity(vel.shape[1:-1])[None]
    jac = np.broadcast_to(np.eye(3), vel.shape[:-1] + (3, 3)).copy()
    for _ in range(steps):
        jac = composition_jac(jac, vel, displacement=True)
        vel = vel + pull_vel(vel, ident + vel, order, bound)
    return vel, jac


def exp_backward(vel, grad, hess=None, inverse=False, steps=8, interpolation="linear", bound="dft",
                 rotate_grad=False):
    """DARTEL-style backward of the exponentiation (_svf.py:189-297)."""
    order = _ORDERS.get(interpolation, interpolation)
    vel = np.array(vel, np.float64, copy=True)
    grad = np.asarray(grad, np.float64)
    ident = identity(vel.shape[1:-1])[None]
    if rotate_grad:
        _, jac = exp_forward_jac(vel, steps, interpolation, bound)
        jac = np.swapaxes(jac, -1, -2)
        grad = (jac @ grad[..., None])[..., 0]
        if hess is not None:
            hess = jhj(jac, hess)
    vel = vel / ((-1 if not inverse else 1) * (2 ** steps))
    jac = grid_jacobian(vel, bound=bound, type="disp")
    ijac = np.linalg.inv(np.swapaxes(2 * np.eye(3) - jac, -1, -2))
    grad = (ijac @ grad[..., None])[..., 0]
    for _ in range(steps):
        det = np.linalg.det(jac)
        jt = np.swapaxes(jac, -1, -2)
        grad0 = grad
        grad = pull_vel(grad, ident + vel, order, bound)
        grad = (jt @ grad[..., None])[..., 0] * det[..., None] + grad0
        if hess is not None:
            hess0 = hess
            hess = pull_vel(hess, ident + vel, order, bound)
            hess = jhj(jt, hess) * det[..., None] + hess0
        jac = composition_jac(jac, vel, displacement=True)
        vel = vel + pull_vel(vel, ident + vel, order, bound)
    if inverse:
        grad = -grad
    grad = grad / (2 ** steps)
    if hess is not None:
        return grad, hess / (2 ** steps)
    return grad
