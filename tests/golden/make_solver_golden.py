"""Write tests/golden/solver_golden.npz: vectors of the reference's own
compiled CPU backend (oracle/_ref) for the displacement-field solver path:
regulariser_grid, relax_grid, fmg_prolongation, fmg_restriction, fmg_grid
(Gauss-Seidel and PCG smoothers), precond_grid, pcg_grid (float64, small 3-D
fields, batch 1).

    python tests/golden/make_solver_golden.py

Notes on the reference (reproduced by the oracle only under its quirk flags):
relax_grid uses the fold start indices of relax_grid_common.cpp:469-470 and
the sign typos of :1154-1158; fmg_grid rescales Z-slices
(fmg_common.cpp:392-489).  Batch 1 only: regulariser_grid offsets the
Hessian of batch n by n * stride(input, 0) (regulariser_grid_common.cpp:814-818).
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref as R  # noqa: E402

PENALTIES = [
    (1e-3, 0.0, 0.0, 0.0, 0.0),
    (1e-4, 0.1, 0.0, 0.0, 0.0),
    (1e-4, 1e-3, 0.2, 0.0, 0.0),
    (1e-4, 1e-3, 0.2, 0.05, 0.2),
    (0.0, 0.1, 0.0, 0.3, 0.1),
]
BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]


def main():
    m = R.load()
    if m is None:
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref -j8")
    rng = np.random.default_rng(4321)
    out = {}
    k = 0
    E = torch.Tensor().double()
    for shape in [(6, 5, 7)]:
        for bname in ("dft", "dct2", "replicate", "dst2"):
            for ip, p in enumerate(PENALTIES):
                for CC in ((0, 1, 3, 6)[ip % 4], 6):
                    vs = [1.0, 1.3, 0.8]
                    x = rng.standard_normal((1, 3) + shape)
                    g = rng.standard_normal((1, 3) + shape)
                    h = None
                    if CC:
                        h = rng.uniform(0.5, 2, (1, CC) + shape)
                        if CC == 6:
                            h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
                    th = torch.from_numpy(h) if CC else E
                    B = R.bounds([bname])
                    reg = m.regulariser_grid(torch.from_numpy(x), E, E, th, *p, vs, B).numpy()
                    key = f"r{k:04d}"
                    out[key + "_meta"] = np.array([BOUNDS.index(bname), CC], np.int64)
                    out[key + "_pen"] = np.array(p + tuple(vs))
                    out[key + "_x"] = x
                    out[key + "_g"] = g
                    if CC:
                        out[key + "_h"] = h
                    out[key + "_reg"] = reg
                    if bname != "zero":
                        rel = m.relax_grid(th, torch.from_numpy(g), torch.from_numpy(x.copy()), E, *p, vs, B, 2).numpy()
                        out[key + "_relax"] = rel
                    k += 1
    # transfers
    t = 0
    for shape in [(7, 6, 5), (3, 3, 3), (2, 2, 2), (5, 1, 4)]:
        for bname in BOUNDS:
            x = rng.standard_normal((1, 3) + shape)
            cs = tuple(1 + (s - 1) // 2 for s in shape)
            xc = rng.standard_normal((1, 3) + cs)
            B = R.bounds([bname])
            key = f"t{t:04d}"
            out[key + "_meta"] = np.array([BOUNDS.index(bname)], np.int64)
            out[key + "_fine"] = x
            out[key + "_coarse"] = xc
            out[key + "_restrict"] = m.fmg_restriction(torch.from_numpy(x), torch.zeros((1, 3) + cs, dtype=torch.float64), B).numpy()
            out[key + "_prolong"] = m.fmg_prolongation(torch.from_numpy(xc), torch.zeros((1, 3) + shape, dtype=torch.float64), B).numpy()
            t += 1
    # fmg
    f = 0
    for shape in [(12, 10, 9)]:
        for bname in ("dft", "dct2"):
            for p in PENALTIES[2:]:
                g = rng.standard_normal((1, 3) + shape)
                h = rng.uniform(0.5, 2, (1, 6) + shape)
                h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
                sol = m.fmg_grid(torch.from_numpy(h), torch.from_numpy(g), E, E, *p, [1., 1., 1.],
                                 R.bounds([bname]), 2, 2, 16, False).numpy()
                key = f"f{f:04d}"
                out[key + "_meta"] = np.array([BOUNDS.index(bname)], np.int64)
                out[key + "_pen"] = np.array(p)
                out[key + "_g"] = g
                out[key + "_h"] = h
                out[key + "_sol"] = sol
                f += 1
    # precond_grid / pcg_grid (tol 0: fixed iterations; tol > 0: early stop)
    c = 0
    for shape in [(6, 5, 7)]:
        for bname in ("dft", "dct2", "replicate", "dst2", "zero", "dct1"):
            for ip, p in enumerate(PENALTIES):
                CC = (0, 1, 3, 6)[(ip + c) % 4]
                vs = [1.0, 1.3, 0.8]
                g = rng.standard_normal((1, 3) + shape)
                x = rng.standard_normal((1, 3) + shape) * 0.1
                h = None
                if CC:
                    h = rng.uniform(0.5, 2, (1, CC) + shape)
                    if CC == 6:
                        h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
                th = torch.from_numpy(h) if CC else E
                B = R.bounds([bname])
                key = f"c{c:04d}"
                out[key + "_meta"] = np.array([BOUNDS.index(bname), CC], np.int64)
                out[key + "_pen"] = np.array(p + tuple(vs))
                out[key + "_g"] = g
                out[key + "_x"] = x
                if CC:
                    out[key + "_h"] = h
                out[key + "_precond"] = m.precond_grid(th, torch.from_numpy(g), E, E, *p, vs, B).numpy()
                out[key + "_pcg"] = m.pcg_grid(th, torch.from_numpy(g), torch.from_numpy(x.copy()), E,
                                               *p, vs, B, 5, 0.0).numpy()
                out[key + "_pcgtol"] = m.pcg_grid(th, torch.from_numpy(g), torch.from_numpy(x.copy()), E,
                                                  *p, vs, B, 50, 1e-3).numpy()
                c += 1
    # fmg_grid with the PCG smoother (use_cg)
    for shape in [(12, 10, 9)]:
        for bname in ("dft", "dct2"):
            for p in PENALTIES[2:]:
                g = rng.standard_normal((1, 3) + shape)
                h = rng.uniform(0.5, 2, (1, 6) + shape)
                h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
                sol = m.fmg_grid(torch.from_numpy(h), torch.from_numpy(g), E, E, *p, [1., 1., 1.],
                                 R.bounds([bname]), 2, 2, 16, True).numpy()
                key = f"f{f:04d}"
                out[key + "_meta"] = np.array([BOUNDS.index(bname)], np.int64)
                out[key + "_pen"] = np.array(p)
                out[key + "_g"] = g
                out[key + "_h"] = h
                out[key + "_solcg"] = sol
                f += 1
    path = os.path.join(HERE, "solver_golden.npz")
    np.savez_compressed(path, **out)
    print(path, k, t, f, c)


if __name__ == "__main__":
    main()
