"""nitorch_amd: MI355X-native spline push/pull resampling and full-multigrid
solver for regularised displacement fields, drop-in for nitorch's compiled
backend (nitorch/_C) and its spatial entry points.

    from nitorch_amd import spatial
    out = spatial.grid_pull(img, grid, interpolation=3, bound='dct2')

The HIP library is built in-tree by `python -m nitorch_amd.build`.
"""
from . import grid  # noqa: F401
from . import spatial  # noqa: F401
from .spatial import (grid_pull, grid_push, grid_count, grid_grad,  # noqa: F401
                      identity_grid, add_identity_grid, add_identity_grid_,
                      sub_identity_grid, sub_identity_grid_)
from .grid import BoundType, InterpolationType  # noqa: F401

__version__ = "0.1.0"
