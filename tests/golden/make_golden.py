"""Write tests/golden/pushpull_golden.npz: input/output vectors of the
reference's own compiled CPU backend (nitorch/_C, built from the reference
sources by oracle/Makefile.ref into oracle/_ref/) for small seeded cases of
every push/pull entry point.

    python tests/golden/make_golden.py

Each case stores the inputs (float64 and float32 variants), the options and
the reference outputs.  The fixtures pin the oracle restatement; the tests
that read them need neither the reference nor a GPU.
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


def make_case(rng, dim, dtype):
    sp = (4, 5, 3)[:dim]
    spo = (3, 4, 2)[:dim]
    inp = rng.standard_normal((2, 2) + sp)
    grid = np.stack(np.meshgrid(*[np.arange(s, dtype=np.float64) for s in spo], indexing="ij"), -1)[None]
    grid = grid * (np.array(sp) - 1) / np.maximum(np.array(spo) - 1, 1)
    grid = grid + 1.3 * rng.standard_normal((2,) + spo + (dim,))
    return inp.astype(dtype), np.ascontiguousarray(grid.astype(dtype))


def main():
    m = R.load()
    if m is None:
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref -j8")
    rng = np.random.default_rng(1234)
    out = {}
    idx = 0
    for dim in (1, 2, 3):
        for order in range(8):
            for bname in BOUNDS:
                # float32 everywhere, float64 on every other bound
                dtypes = (np.float32, np.float64) if BOUNDS.index(bname) % 2 == 0 else (np.float32,)
                for dtype in dtypes:
                    extrap = int(rng.integers(0, 2))
                    inp, grid = make_case(rng, dim, dtype)
                    ti = torch.from_numpy(inp).requires_grad_()
                    tg = torch.from_numpy(grid).requires_grad_()
                    B, I = R.bounds([bname]), R.orders([order])
                    shape = list(inp.shape[2:])
                    src_push = rng.standard_normal((2, 2) + grid.shape[1:-1]).astype(dtype)
                    gpull = rng.standard_normal((2, 2) + grid.shape[1:-1]).astype(dtype)
                    gcount = rng.standard_normal((2, 1) + inp.shape[2:]).astype(dtype)
                    with torch.no_grad():
                        pull = m.grid_pull(ti, tg, B, I, extrap).numpy()
                        # 1-D grid_grad without extrapolation writes one element past
                        # its output in the reference (pushpull_common.cpp:1037-1043,
                        # out_sK of a [N, C, X, 1] tensor): not called there.
                        sgrad = None if (dim == 1 and extrap == 0) else m.grid_grad(ti, tg, B, I, extrap).numpy()
                        push = m.grid_push(torch.from_numpy(src_push), tg, shape, B, I, extrap).numpy()
                        count = m.grid_count(tg, shape, B, I, extrap).numpy()
                    # the same 1-D out-of-bound zeroing writes past the grid gradient
                    # (pushpull_common.cpp:1049-1054): grid gradients skipped there.
                    bad1d = dim == 1 and extrap == 0
                    tgb = torch.from_numpy(grid).requires_grad_(not bad1d)
                    pbw = m.grid_pull_backward(torch.from_numpy(gpull), ti, tgb, B, I, extrap)
                    tsp = torch.from_numpy(src_push).requires_grad_()
                    ubw = m.grid_push_backward(torch.from_numpy(inp), tsp, tgb, B, I, extrap)
                    cbw = None if bad1d else m.grid_count_backward(torch.from_numpy(gcount), tgb, B, I, extrap)
                    key = f"c{idx:04d}"
                    out[key + "_meta"] = np.array([dim, order, BOUNDS.index(bname), extrap,
                                                   0 if dtype == np.float32 else 1], np.int64)
                    out[key + "_inp"] = inp
                    out[key + "_grid"] = grid
                    out[key + "_srcpush"] = src_push
                    out[key + "_gpull"] = gpull
                    out[key + "_gcount"] = gcount
                    out[key + "_pull"] = pull
                    if sgrad is not None:
                        out[key + "_sgrad"] = sgrad
                    out[key + "_push"] = push
                    out[key + "_count"] = count
                    out[key + "_pbw_in"] = pbw[0].detach().numpy()
                    out[key + "_ubw_in"] = ubw[0].detach().numpy()
                    if not bad1d:
                        out[key + "_pbw_grid"] = pbw[1].detach().numpy()
                        out[key + "_ubw_grid"] = ubw[1].detach().numpy()
                        out[key + "_cbw_grid"] = cbw.detach().numpy()
                    idx += 1
    path = os.path.join(HERE, "pushpull_golden.npz")
    np.savez_compressed(path, **out)
    print(path, idx, "cases")


if __name__ == "__main__":
    main()
