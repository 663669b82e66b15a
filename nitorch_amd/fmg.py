"""Full multigrid whose transfers are grid_pull / grid_push.

Restates the reference's non-compiled multigrid for displacement fields
(nitorch/spatial/_solvers.py: `prolong` / `restrict` :147-222, `make_grid`
:224-236, `_FMG.allocate` :294-342, `_FMG.solve` :348-391, `_GridFMG`
rescaling :700-740) on this package's kernels:

* prolongation = `grid_pull` (quadratic B-spline, extrapolate) at the fine
  voxel centres expressed in coarse voxels; restriction = `grid_push`
  (linear B-spline, extrapolate) of the fine field onto the coarse grid,
  times prod(n_coarse) / prod(n_fine);
* displacement components are rescaled by the per-axis size ratios, the
  Hessian's diagonal by their squares and its off-diagonal by their pairwise
  products, as `_GridFMG` does (the compiled reference rescales z-slices
  instead, fmg_common.cpp:389-490);
* each level's smoother is the compiled one: block-Jacobi PCG
  (`optim='cg'`, pcg_common.cpp:38-85, the recurrence of
  nitorch/core/optim.py:cg with tolerance 0) or multi-colour Gauss-Seidel
  (`optim='relax'`), with the voxel size scaled by n_0 / n_level.

Every array operation is a HIP kernel of libnitorch_amd.so or a torch
elementwise op on the GPU tensor; there is no CPU path.
Tensors are channel-first: gradient / solution [N, 3, X, Y, Z], Hessian
[N, CC, X, Y, Z] with CC in {1, 3, 6} (h00 h11 h22 h01 h02 h12).
"""
import math

import torch

from . import solve as SV
from . import spatial as SP


def make_grid(highshape, lowshape, device=None, dtype=None):
    """Sampling grid of a fine lattice in coarse voxel coordinates
    (nitorch/spatial/_solvers.py:224-236): [*highshape, 3]."""
    axes = []
    for lo, hi in zip(lowshape, highshape):
        shift = 0.5 * (lo / hi - 1)
        axes.append(torch.arange(0., hi, device=device, dtype=dtype).mul_(lo / hi).add_(shift))
    return torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)


def prolong(x, shape, bound="dct2", order=2, grid=None):
    """Coarse [N, C, *spatial] -> fine [N, C, *shape] by grid_pull
    (nitorch/spatial/_solvers.py:147-181)."""
    if grid is None:
        grid = make_grid(shape, x.shape[2:], x.device, x.dtype)
    return SP.grid_pull(x, grid, interpolation=order, bound=bound, extrapolate=True)


def restrict(x, shape, bound="dct2", order=1, grid=None):
    """Fine [N, C, *spatial] -> coarse [N, C, *shape] by grid_push, normalised
    by the change of resolution (nitorch/spatial/_solvers.py:184-222)."""
    if grid is None:
        grid = make_grid(x.shape[2:], shape, x.device, x.dtype)
    y = SP.grid_push(x, grid, shape, interpolation=order, bound=bound, extrapolate=True)
    return y.mul_(math.prod(shape) / math.prod(x.shape[2:]))


def _axis_scale(a, b, like):
    return torch.tensor([p / q for p, q in zip(a, b)], dtype=like.dtype, device=like.device)


class GridFMG:
    """Multigrid solver of (H + L) x = g for displacement fields
    (`_GridFMG`, nitorch/spatial/_solvers.py:686-740, with `_FMG.solve`)."""

    def __init__(self, absolute=0, membrane=0, bending=0, lame=0, factor=1, voxel_size=1,
                 bound="dft", nb_cycles=2, nb_iter=2, max_levels=16, optim="cg"):
        if optim == "relaks":
            optim = "relax"
        if optim not in ("cg", "relax"):
            raise ValueError(f"optim must be 'cg' or 'relax' (got {optim!r})")
        self.penalty = SV._penalty(absolute, membrane, bending, lame, factor)
        self.voxel_size = list(SV._voxel(voxel_size))
        self.bound = bound
        self.nb_cycles = int(nb_cycles)
        self.nb_iter = int(nb_iter)
        self.max_levels = int(max_levels)
        self.optim = optim

    # -- transfers with the displacement-field rescaling (_solvers.py:700-740)
    def prolong_g(self, x, grid, shape):
        scale = _axis_scale(shape, x.shape[2:], x)
        return prolong(x, shape, self.bound, grid=grid).mul_(scale.view(1, -1, 1, 1, 1))

    def restrict_g(self, x, grid, shape):
        scale = _axis_scale(x.shape[2:], shape, x)
        return restrict(x, shape, self.bound, grid=grid).mul_(scale.view(1, -1, 1, 1, 1))

    def restrict_h(self, h, grid, shape):
        scale = _axis_scale(h.shape[2:], shape, h)
        h = restrict(h, shape, self.bound, grid=grid)
        h[:, :3].mul_(scale.square().view(1, -1, 1, 1, 1)[:, :h.shape[1]])
        if h.shape[1] == 6:
            h[:, 3].mul_(scale[0] * scale[1])
            h[:, 4].mul_(scale[0] * scale[2])
            h[:, 5].mul_(scale[1] * scale[2])
        return h

    # -- pyramids (_FMG.allocate, _solvers.py:294-342)
    def allocate(self, hessian, gradient, init=None):
        spatial = list(gradient.shape[2:])
        self.pyrn = [spatial]
        self.pyrv = [[1.0, 1.0, 1.0]]
        self.pyrt = []
        self.pyrh = [hessian]
        self.pyrg = [gradient]
        zero_init = init is None
        self.pyrx = [torch.zeros_like(gradient) if zero_init else init]
        for _ in range(1, self.max_levels + 1):
            sp1 = [math.ceil(s / 2) for s in self.pyrn[-1]]
            if all(s == 1 for s in sp1):
                break
            self.pyrt.append(make_grid(self.pyrn[-1], sp1, gradient.device, gradient.dtype))
            self.pyrn.append(sp1)
            self.pyrv.append([n0 / n for n, n0 in zip(sp1, self.pyrn[0])])
            self.pyrh.append(None if hessian is None else self.restrict_h(self.pyrh[-1], self.pyrt[-1], sp1))
            self.pyrg.append(self.restrict_g(self.pyrg[-1], self.pyrt[-1], sp1))
            self.pyrx.append(torch.zeros_like(self.pyrg[-1]) if zero_init
                             else self.restrict_g(self.pyrx[-1], self.pyrt[-1], sp1))

    def _vx(self, n):
        return [v0 * v for v0, v in zip(self.voxel_size, self.pyrv[n])]

    def smooth(self, n, g, x):
        """nb_iter steps of the level-n smoother on (H_n + L_n) x = g (in place)."""
        h, vx = self.pyrh[n], self._vx(n)
        if self.optim == "cg":
            SV.pcg_grid_cf(h, g, x, self.penalty, vx, self.bound, self.nb_iter, 0.0)
        else:
            SV.relax_grid_cf(g, x, h, self.penalty, vx, self.bound, self.nb_iter)
        return x

    def residual(self, n, g, x):
        """g - (H_n + L_n) x."""
        return g - SV.regulariser_grid_cf(x, self.pyrh[n], self.penalty, self._vx(n), self.bound)

    # -- _FMG.solve (_solvers.py:348-391)
    def solve(self, hessian, gradient, init=None):
        g_ = gradient.contiguous()
        h_ = None if hessian is None else hessian.contiguous()
        self.allocate(h_, g_, None if init is None else init.contiguous().clone())
        g, x, d, t = self.pyrg, self.pyrx, self.pyrn, self.pyrt
        nl = len(self.pyrn)
        self.smooth(nl - 1, g[-1], x[-1])
        for n_base in reversed(range(nl - 1)):
            x[n_base] = self.prolong_g(x[n_base + 1], t[n_base], d[n_base])
            for _ in range(self.nb_cycles):
                for n in range(n_base, nl - 1):
                    self.smooth(n, g[n], x[n])
                    g[n + 1] = self.restrict_g(self.residual(n, g[n], x[n]), t[n], d[n + 1])
                    x[n + 1].zero_()
                self.smooth(nl - 1, g[-1], x[-1])
                for n in reversed(range(n_base, nl - 1)):
                    x[n] += self.prolong_g(x[n + 1], t[n], d[n])
                    self.smooth(n, g[n], x[n])
        out = x[0]
        self.pyrh = self.pyrg = self.pyrx = self.pyrt = []
        return out


def solve_grid_fmg_pushpull(hessian, gradient, absolute=0, membrane=0, bending=0, lame=0, factor=1,
                            voxel_size=1, bound="dft", optim="cg", nb_cycles=2, nb_iter=2,
                            max_levels=16):
    """`solve_grid_fmg` through the pull/push multigrid, on the reference's
    channel-last layout: hessian (N, *spatial, CC), gradient (N, *spatial, 3)
    -> solution (N, *spatial, 3)."""
    g = gradient.movedim(-1, 1).contiguous()
    h = None if hessian is None else hessian.movedim(-1, 1).contiguous()
    F = GridFMG(absolute, membrane, bending, lame, factor, voxel_size, bound, nb_cycles, nb_iter,
                max_levels, optim)
    return F.solve(h, g).movedim(1, -1)
