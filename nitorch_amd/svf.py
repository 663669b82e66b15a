"""Exponentiation of stationary velocity fields by scaling and squaring.

Restates nitorch/spatial/_svf.py (`exp` :11-42, `exp_forward` :45-99,
`_pull_vel` :323-361, `_push_vel` :364-402) on this package's grid_pull /
grid_push: the velocity is divided by 2**steps and composed with itself
`steps` times, v <- v + v o (id + v), each composition one order-`order`
pull of the 3-channel field (extrapolate, dft by default).  Gradients with
respect to the velocity flow through the pull's backward kernels (a push of
the upstream gradient and the grid gradient), as the reference's autodiff
path (`anagrad=False`).

Fields are channel-last like the reference: ([batch], *spatial, dim).
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


def exp_forward(vel, inverse=False, steps=8, interpolation="linear", bound="dft", displacement=False):
    """Scaling and squaring (nitorch/spatial/_svf.py:45-99, without the
    Jacobian accumulation): returns the transformation (or, with
    `displacement`, the displacement) exp(vel) of shape vel.shape."""
    vel = -vel if inverse else vel.clone()
    dim = vel.shape[-1]
    spatial = vel.shape[-1 - dim:-1]
    ident = SP.identity_grid(spatial, dtype=vel.dtype, device=vel.device)
    iadd = (lambda x, y: x.add(y)) if vel.requires_grad else (lambda x, y: x.add_(y))
    vel = vel / (2 ** steps)
    for _ in range(steps):
        vel = iadd(vel, _pull_vel(vel, ident + vel, interpolation=interpolation, bound=bound))
    if not displacement:
        vel = vel + ident
    return vel


def exp(vel, inverse=False, steps=8, interpolation="linear", bound="dft", displacement=False,
        anagrad=False):
    """Exponentiate a stationary velocity field (nitorch/spatial/_svf.py:11-42).
    Gradients use autodiff through the HIP pull kernels; `anagrad=True`
    (the reference's hand-written DARTEL backward) is not provided."""
    if anagrad:
        raise NotImplementedError("anagrad=True is not provided; autodiff through grid_pull is used")
    return exp_forward(vel, inverse, steps, interpolation, bound, displacement)
