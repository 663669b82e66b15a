"""Write tests/golden/pushpull_golden_edge.npz: float32 vectors of the
reference's own compiled CPU backend (nitorch/_C built from the reference
sources by oracle/Makefile.ref into oracle/_ref/) on grids whose coordinates
sit exactly on the out-of-bound edges.

For float data the reference decides "in bounds" in float arithmetic,
`x >= -tol && x < (float)(n - 1) + tol` with `tol = (float)TINY` (or
`(float)(0.5 + TINY)` for extrapolate=2), bounds_common.h:168-171 and
pushpull_common.cpp:900.  A float32 coordinate such as float32(44.05) on a
size-45 axis is out of bounds in that test and in bounds in an FP64 one, so
these vectors pin the oracle's `oob_f32` option.  Every edge value is
written with its two float neighbours; the other coordinates of each sample
are interior.

    python tests/golden/make_golden_edge.py
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


def edge_values(n, extrapolate):
    """float32 edge coordinates of an axis of size n and their neighbours."""
    tol = 0.5 + 5e-2 if extrapolate == 2 else 5e-2
    vals = []
    for e in (np.float32(-tol), np.float32(n - 1 + tol)):
        vals += [np.nextafter(e, np.float32(-np.inf)), e, np.nextafter(e, np.float32(np.inf))]
    return np.array(vals, np.float32)


def edge_grid(rng, sp, nsamp, extrapolate):
    """[1, nsamp, dim] (1-D list of samples): each sample puts one axis on
    an edge value, the others uniformly inside."""
    dim = len(sp)
    g = np.empty((1, nsamp, dim), np.float32)
    for s in range(nsamp):
        ax = s % dim
        for d in range(dim):
            if d == ax:
                ev = edge_values(sp[d], extrapolate)
                g[0, s, d] = ev[(s // dim) % len(ev)]
            else:
                g[0, s, d] = np.float32(rng.uniform(0.2, sp[d] - 1.2))
    return g


def main():
    m = R.load()
    if m is None:
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref -j8")
    rng = np.random.default_rng(8642)
    out = {}
    idx = 0
    # the lower edge float32(-tol) is out of bounds in FP64 and in bounds in
    # float for every size; the upper one differs for some sizes (3, 4, 17,
    # 20, 33, 45, 46) and not for others (5, 6)
    for sp in ((45, 6, 5), (5, 33, 4), (4, 5, 46), (17, 20, 3)):
        for order in (0, 1, 3):
            for bname in ("dct2", "zero", "dft"):
                for extrap in (0, 2):
                    B, I = R.bounds([bname]), R.orders([order])
                    nsamp = 3 * 6 * 4
                    grid = edge_grid(rng, sp, nsamp, extrap)
                    # the reference takes [N, *spatial_out, dim]: one row of samples
                    grid = grid.reshape(1, nsamp, 1, 1, 3)
                    inp = rng.standard_normal((1, 2) + sp).astype(np.float32)
                    src_push = rng.standard_normal((1, 2, nsamp, 1, 1)).astype(np.float32)
                    key = f"d{idx:04d}"
                    out[key + "_meta"] = np.array([3, order, BOUNDS.index(bname), extrap], np.int64)
                    out[key + "_inp"] = inp
                    out[key + "_grid"] = grid
                    out[key + "_srcpush"] = src_push
                    tg = torch.from_numpy(grid)
                    with torch.no_grad():
                        out[key + "_pull"] = m.grid_pull(torch.from_numpy(inp), tg, B, I, extrap).numpy()
                        out[key + "_push"] = m.grid_push(torch.from_numpy(src_push), tg, list(sp), B, I,
                                                         extrap).numpy()
                        out[key + "_count"] = m.grid_count(tg, list(sp), B, I, extrap).numpy()
                    idx += 1
    path = os.path.join(HERE, "pushpull_golden_edge.npz")
    np.savez_compressed(path, **out)
    print(path, idx, "cases")


if __name__ == "__main__":
    main()
