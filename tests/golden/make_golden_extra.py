"""Write tests/golden/pushpull_golden_extra.npz: vectors of the reference's own
compiled CPU backend (nitorch/_C built from the reference sources by
oracle/Makefile.ref into oracle/_ref/) for the two branches that the first
fixture set (make_golden.py) leaves open:

  * `extrapolate=2`: the widened out-of-bound tolerance 0.5 + TINY of
    pushpull_common.cpp:900 (0.55 here and in the reference's float build),
    on grids whose coordinates fill the band between the 0.05 and 0.55
    tolerances, for pull / push / count / grad and the pull / push / count
    backward passes;
  * the input branch of `grid_grad_backward` (pushpull.cpp:303-327, the
    adjoint of grid_grad with respect to its input), for every extrapolate
    mode.

    python tests/golden/make_golden_extra.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref as R  # noqa: E402

BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]


def band_grid(rng, sp, spo, dim, dtype):
    """Coordinates uniform over [-0.8, n - 1 + 0.8] per axis: a fixed share
    of them falls between the 0.05 and 0.55 tolerances of either edge."""
    lo = -0.8
    hi = np.array(sp, np.float64) - 1 + 0.8
    g = lo + rng.random((1,) + spo + (dim,)) * (hi - lo)
    return np.ascontiguousarray(g.astype(dtype))


def main():
    m = R.load()
    if m is None:
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref -j8")
    rng = np.random.default_rng(4321)
    out = {}
    idx = 0
    for dim in (1, 2, 3):
        sp = (5, 6, 4)[:dim]
        spo = (4, 3, 5)[:dim]
        for order in range(8):
            for bname in BOUNDS:
                # float32 everywhere, float64 on every other bound
                dtypes = (np.float32, np.float64) if BOUNDS.index(bname) % 2 == 0 else (np.float32,)
                for dtype in dtypes:
                    B, I = R.bounds([bname]), R.orders([order])
                    inp = rng.standard_normal((1, 2) + sp).astype(dtype)
                    grid = band_grid(rng, sp, spo, dim, dtype)
                    src_push = rng.standard_normal((1, 2) + spo).astype(dtype)
                    gpull = rng.standard_normal((1, 2) + spo).astype(dtype)
                    gcount = rng.standard_normal((1, 1) + sp).astype(dtype)
                    ggrad = rng.standard_normal((1, 2) + spo + (dim,)).astype(dtype)
                    ti = torch.from_numpy(inp)
                    tg = torch.from_numpy(grid)
                    key = f"e{idx:04d}"
                    out[key + "_meta"] = np.array([dim, order, BOUNDS.index(bname), 2,
                                                   0 if dtype == np.float32 else 1], np.int64)
                    out[key + "_inp"] = inp
                    out[key + "_grid"] = grid
                    out[key + "_srcpush"] = src_push
                    out[key + "_gpull"] = gpull
                    out[key + "_gcount"] = gcount
                    out[key + "_ggrad"] = ggrad
                    with torch.no_grad():
                        out[key + "_pull"] = m.grid_pull(ti, tg, B, I, 2).numpy()
                        out[key + "_push"] = m.grid_push(torch.from_numpy(src_push), tg, list(sp), B, I, 2).numpy()
                        out[key + "_count"] = m.grid_count(tg, list(sp), B, I, 2).numpy()
                    # 1-D grid_grad / grid gradients zero their out-of-bound
                    # outputs one element past the end in the reference
                    # (pushpull_common.cpp:1037-1054): not called in 1-D.
                    if dim > 1:
                        with torch.no_grad():
                            out[key + "_sgrad"] = m.grid_grad(ti, tg, B, I, 2).numpy()
                    tgb = torch.from_numpy(grid).requires_grad_(dim > 1)
                    tin = torch.from_numpy(inp).requires_grad_()
                    pbw = m.grid_pull_backward(torch.from_numpy(gpull), tin, tgb, B, I, 2)
                    out[key + "_pbw_in"] = pbw[0].detach().numpy()
                    tsp = torch.from_numpy(src_push).requires_grad_()
                    ubw = m.grid_push_backward(torch.from_numpy(inp), tsp, tgb, B, I, 2)
                    out[key + "_ubw_in"] = ubw[0].detach().numpy()
                    if dim > 1:
                        out[key + "_pbw_grid"] = pbw[1].detach().numpy()
                        out[key + "_ubw_grid"] = ubw[1].detach().numpy()
                        out[key + "_cbw_grid"] = m.grid_count_backward(torch.from_numpy(gcount), tgb, B, I,
                                                                       2).detach().numpy()
                    # grid_grad_backward's input branch at every extrapolate
                    # mode.  The grid requires grad too: without it the
                    # reference skips the spline derivatives
                    # (pushpull_common.cpp:1137, `if (do_grad || do_sgrad)`)
                    # that this branch multiplies (:1237-1244), and reads them
                    # uninitialised.  Its grid output is discarded (it reads
                    # uninitialised second derivatives when trgt_sK <= 1).
                    for ex in (0, 1, 2):
                        if dim == 1 and ex != 1:
                            continue   # same 1-D out-of-bound overrun as above
                        tgr = torch.from_numpy(grid).requires_grad_()
                        gbw = m.grid_grad_backward(torch.from_numpy(ggrad), tin, tgr, B, I, ex)
                        out[key + f"_gbw_in{ex}"] = gbw[0].detach().numpy()
                    idx += 1
    path = os.path.join(HERE, "pushpull_golden_extra.npz")
    np.savez_compressed(path, **out)
    print(path, idx, "cases")


if __name__ == "__main__":
    main()
