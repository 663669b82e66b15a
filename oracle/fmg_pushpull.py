"""ORACLE — test infrastructure only.

FP64 numpy restatement of the reference's non-compiled multigrid for
displacement fields, whose transfers are pull and push
(nitorch/spatial/_solvers.py: prolong :147-181, restrict :184-222,
make_grid :224-236, _FMG.allocate :294-342, _FMG.solve :348-391, _GridFMG
rescaling :700-740), built from the oracle's own pull / push
(`pushpull_oracle.c`) and level operators (`solver_oracle.c`): block-Jacobi
PCG (`optim='cg'`) or the wrap-safe multi-colour Gauss-Seidel
(`optim='relax'`) as the level smoother, (H + L) x as the level operator.
Arrays are channel-first: g, x [N, 3, X, Y, Z]; H [N, CC, X, Y, Z].
"""
import math

import numpy as np

from . import pushpull as PP
from . import solver as S


def make_grid(highshape, lowshape):
    """[*highshape, 3] coordinates of the fine voxels in coarse voxels
    (_solvers.py:224-236)."""
    axes = [np.arange(hi, dtype=np.float64) * (lo / hi) + 0.5 * (lo / hi - 1)
            for lo, hi in zip(lowshape, highshape)]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1)


def _tile(grid, N):
    return np.ascontiguousarray(np.broadcast_to(grid[None], (N,) + grid.shape))


def prolong(x, shape, bound, grid=None):
    """Quadratic pull, extrapolated (_solvers.py:147-181)."""
    if grid is None:
        grid = make_grid(shape, x.shape[2:])
    return PP.grid_pull(np.asarray(x, np.float64), _tile(grid, x.shape[0]), bound, 2, 1)


def restrict(x, shape, bound, grid=None):
    """Linear push, extrapolated, times prod(coarse) / prod(fine)
    (_solvers.py:184-222)."""
    if grid is None:
        grid = make_grid(x.shape[2:], shape)
    y = PP.grid_push(np.asarray(x, np.float64), _tile(grid, x.shape[0]), tuple(shape), bound, 1, 1)
    return y * (math.prod(shape) / math.prod(x.shape[2:]))


def _scale(a, b):
    return np.array([p / q for p, q in zip(a, b)], np.float64)


def fmg(hessian, g, absolute=0, membrane=0, bending=0, lame=0, factor=1, voxel_size=1,
        bound="dft", nb_cycles=2, nb_iter=2, max_levels=16, optim="cg"):
    """Solution of (H + L) x = g by `_FMG.solve` with pull/push transfers."""
    lame = list(np.atleast_1d(lame).astype(float))
    lame = (lame + [lame[-1]])[:2]
    pen = dict(absolute=absolute * factor, membrane=membrane * factor, bending=bending * factor,
               lame=(lame[0] * factor, lame[1] * factor))
    vs = list(np.atleast_1d(voxel_size).astype(float))
    vs = (vs + [vs[-1]] * 3)[:3]
    g = np.asarray(g, np.float64)
    h = None if hessian is None else np.asarray(hessian, np.float64)

    def restrict_g(x, grid, shape):
        return restrict(x, shape, bound, grid) * _scale(x.shape[2:], shape)[None, :, None, None, None]

    def prolong_g(x, grid, shape):
        return prolong(x, shape, bound, grid) * _scale(shape, x.shape[2:])[None, :, None, None, None]

    def restrict_h(hh, grid, shape):
        s = _scale(hh.shape[2:], shape)
        out = restrict(hh, shape, bound, grid)
        out[:, :3] *= (s * s)[None, :out.shape[1], None, None, None]
        if out.shape[1] == 6:
            out[:, 3] *= s[0] * s[1]
            out[:, 4] *= s[0] * s[2]
            out[:, 5] *= s[1] * s[2]
        return out

    # pyramids
    pyrn, pyrv, pyrt = [list(g.shape[2:])], [[1.0] * 3], []
    pyrh, pyrg, pyrx = [h], [g], [np.zeros_like(g)]
    for _ in range(max_levels):
        sp1 = [math.ceil(s / 2) for s in pyrn[-1]]
        if all(s == 1 for s in sp1):
            break
        pyrt.append(make_grid(pyrn[-1], sp1))
        pyrn.append(sp1)
        pyrv.append([n0 / n for n, n0 in zip(sp1, pyrn[0])])
        pyrh.append(None if h is None else restrict_h(pyrh[-1], pyrt[-1], sp1))
        pyrg.append(restrict_g(pyrg[-1], pyrt[-1], sp1))
        pyrx.append(np.zeros_like(pyrg[-1]))

    def vx(n):
        return [v0 * v for v0, v in zip(vs, pyrv[n])]

    def smooth(n, gg, xx):
        if optim == "cg":
            return S.pcg_grid(pyrh[n], gg, xx, voxel_size=vx(n), bound=bound, nb_iter=nb_iter,
                              tol=0.0, **pen)[0]
        return S.relax_grid(gg, pyrh[n], xx, voxel_size=vx(n), bound=bound, nb_iter=nb_iter,
                            product_order=True, **pen)

    def residual(n, gg, xx):
        return gg - S.regulariser_grid(xx, pyrh[n], voxel_size=vx(n), bound=bound, **pen)

    gp, x, d, t = pyrg, pyrx, pyrn, pyrt
    nl = len(pyrn)
    x[-1] = smooth(nl - 1, gp[-1], x[-1])
    for n_base in reversed(range(nl - 1)):
        x[n_base] = prolong_g(x[n_base + 1], t[n_base], d[n_base])
        for _ in range(nb_cycles):
            for n in range(n_base, nl - 1):
                x[n] = smooth(n, gp[n], x[n])
                gp[n + 1] = restrict_g(residual(n, gp[n], x[n]), t[n], d[n + 1])
                x[n + 1] = np.zeros_like(x[n + 1])
            x[-1] = smooth(nl - 1, gp[-1], x[-1])
            for n in reversed(range(n_base, nl - 1)):
                x[n] = x[n] + prolong_g(x[n + 1], t[n], d[n])
                x[n] = smooth(n, gp[n], x[n])
    return x[0]
