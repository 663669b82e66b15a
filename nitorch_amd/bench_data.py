"""Seeded synthetic inputs for the benchmark workloads.

All random fields are generated on the CPU (torch.Generator, float64),
smoothed in float64 by a separable Gaussian applied in the Fourier domain
(periodic), then cast to float32, so that CPU and GPU paths see
bit-identical coordinates.

Workloads (voxel units):
  pull_96        1x1x96^3 image U[0,1); grid = identity + N(0, 4^2) smoothed
                 with sigma = 6; order 1, dct2, no extrapolation.
  pushpull_192   4x8x192^3 image U[0,1); grid = identity + N(0, 8^2) smoothed
                 with sigma = 10, then 5% of the grid points moved out of
                 bounds by +-(size) along one random axis.
  backward_128   2x3x128x160x128 image, same grid recipe; order 3, dft;
                 upstream gradient N(0, 1).
  fmg_160        Gauss-Newton system on a 160x192x160x3 displacement field:
                 gradient = N(0,1) smoothed (sigma = 2), Hessian 6 channels,
                 diagonal U[0.5, 2], off-diagonal U[-0.2, 0.2].

A unit-sum Gaussian shrinks white noise by ~(4 pi sigma^2)^(-3/4), so the
realised displacement RMS is far below the nominal std; `rescale=True`
rescales the smoothed field back to the nominal std (reported variant).
"""
import math

import torch


def _gauss_fft(shape, sigma, dtype=torch.float64):
    """Fourier transfer function of a unit-sum Gaussian on a periodic grid."""
    out = None
    for i, n in enumerate(shape):
        f = torch.fft.fftfreq(n, dtype=dtype) if i < len(shape) - 1 else torch.fft.rfftfreq(n, dtype=dtype)
        h = torch.exp(-2 * (math.pi * sigma * f) ** 2)
        view = [1] * len(shape)
        view[i] = h.numel()
        h = h.view(view)
        out = h if out is None else out * h
    return out


def smooth(field, sigma):
    """Gaussian smoothing (float64, periodic) over the last 3 dims."""
    spatial = field.shape[-3:]
    H = _gauss_fft(spatial, sigma)
    F = torch.fft.rfftn(field, s=spatial, dim=(-3, -2, -1))
    return torch.fft.irfftn(F * H, s=spatial, dim=(-3, -2, -1))


def identity(shape, dtype=torch.float64):
    mesh = torch.meshgrid(*[torch.arange(s, dtype=dtype) for s in shape], indexing="ij")
    return torch.stack(mesh, -1)


def smooth_displacement(gen, batch, shape, std, sigma, rescale=False):
    noise = torch.randn((batch, 3) + tuple(shape), generator=gen, dtype=torch.float64) * std
    disp = smooth(noise, sigma)
    if rescale:
        disp = disp * (std / disp.std())
    return disp.movedim(1, -1)          # [B, *shape, 3]


def config1_inputs(seed=1, rescale=False):
    """pull_96: returns (input [1,1,96,96,96], grid [1,96,96,96,3]) float32 CPU."""
    gen = torch.Generator().manual_seed(seed)
    shape = (96, 96, 96)
    x = torch.rand((1, 1) + shape, generator=gen, dtype=torch.float64)
    disp = smooth_displacement(gen, 1, shape, 4.0, 6.0, rescale)
    grid = identity(shape) + disp
    return x.float(), grid.float()


def grid_recipe(gen, batch, shape, std=8.0, sigma=10.0, oob_frac=0.05, rescale=False):
    disp = smooth_displacement(gen, batch, shape, std, sigma, rescale)
    grid = identity(shape) + disp
    nvox = grid[..., 0].numel()
    k = int(round(oob_frac * nvox))
    flat = grid.reshape(-1, 3)
    idx = torch.randperm(nvox, generator=gen)[:k]
    axis = torch.randint(0, 3, (k,), generator=gen)
    sign = torch.randint(0, 2, (k,), generator=gen).to(torch.float64) * 2 - 1
    sizes = torch.tensor(shape, dtype=torch.float64)
    flat[idx, axis] += sign * sizes[axis]
    return flat.view(grid.shape)


def config2_inputs(seed=2, batch=4, channels=8, shape=(192, 192, 192), rescale=False):
    """pushpull_192: (input [4,8,192^3], grid [4,192^3,3]) float32 CPU."""
    gen = torch.Generator().manual_seed(seed)
    x = torch.rand((batch, channels) + tuple(shape), generator=gen, dtype=torch.float32)
    grid = grid_recipe(gen, batch, shape, rescale=rescale)
    return x, grid.float()


def config3_inputs(seed=3, batch=2, channels=3, shape=(128, 160, 128), rescale=False):
    """backward_128: (input, grid, upstream grad) float32 CPU."""
    gen = torch.Generator().manual_seed(seed)
    x = torch.rand((batch, channels) + tuple(shape), generator=gen, dtype=torch.float32)
    grid = grid_recipe(gen, batch, shape, rescale=rescale).float()
    grad = torch.randn((batch, channels) + tuple(shape), generator=gen, dtype=torch.float32)
    return x, grid, grad


def config4_inputs(seed=4, shape=(160, 192, 160), sigma_g=2.0):
    """fmg_160: (hessian [1,*shape,6], gradient [1,*shape,3]) float32 CPU, channel-last
    like nitorch.spatial.solve_grid_fmg."""
    gen = torch.Generator().manual_seed(seed)
    g = smooth(torch.randn((1, 3) + tuple(shape), generator=gen, dtype=torch.float64), sigma_g)
    g = g / g.std()
    hd = torch.rand((1, 3) + tuple(shape), generator=gen, dtype=torch.float64) * 1.5 + 0.5
    ho = torch.rand((1, 3) + tuple(shape), generator=gen, dtype=torch.float64) * 0.4 - 0.2
    h = torch.cat([hd, ho], 1)
    return h.movedim(1, -1).float().contiguous(), g.movedim(1, -1).float().contiguous()
