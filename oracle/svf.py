"""ORACLE — test infrastructure only.

FP64 numpy restatement of scaling and squaring of a stationary velocity
field (nitorch/spatial/_svf.py: exp_forward :45-99, _pull_vel :323-361) on
the oracle's pull (`pushpull_oracle.c`, extrapolated).
Fields are channel-last: vel [N, *spatial, 3].
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
