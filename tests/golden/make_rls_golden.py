"""Write tests/golden/solver_rls_golden.npz: vectors of the reference's own
compiled CPU backend (oracle/_ref) for the solver path with voxel-wise RLS
weights (absolute / membrane energies, the only ones the reference weights):
regulariser_grid, relax_grid, precond_grid, pcg_grid and fmg_grid (both
smoothers), float64, small 3-D fields, batch 1.

    python tests/golden/make_rls_golden.py

Reference behaviour that the vectors contain and the oracle / product do not
reproduce: relax_grid with an absolute-only penalty multiplies the weight by
`absolute` twice on the diagonal (relax_grid_common.cpp:1500-1518), so no
absolute-only relax vector is written; precond_grid reads the weights of
batch 0 for every batch element (precond_grid_common.cpp:810), hence batch 1.
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref as R  # noqa: E402

PENALTIES = [
    (1e-3, 0.0, 0.0, 0.0, 0.0),     # absolute only
    (1e-4, 0.1, 0.0, 0.0, 0.0),     # absolute + membrane
    (0.0, 0.05, 0.0, 0.0, 0.0),     # membrane only
    (0.02, 0.3, 0.0, 0.0, 0.0),
]
BOUNDS = ["replicate", "dct1", "dct2", "dst1", "dst2", "dft", "zero"]


def main():
    m = R.load()
    if m is None:
        raise SystemExit("build the reference first: make -f oracle/Makefile.ref -j8")
    rng = np.random.default_rng(8642)
    out = {}
    E = torch.Tensor().double()
    k = 0
    shape = (6, 5, 7)
    for bname in ("dft", "dct2", "replicate", "dst2", "zero"):
        for ip, p in enumerate(PENALTIES):
            CC = (0, 1, 3, 6)[(ip + k) % 4]
            vs = [1.0, 1.3, 0.8]
            x = rng.standard_normal((1, 3) + shape)
            g = rng.standard_normal((1, 3) + shape)
            w = rng.uniform(0.2, 2.0, (1, 1) + shape)
            h = None
            if CC:
                h = rng.uniform(0.5, 2, (1, CC) + shape)
                if CC == 6:
                    h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
            th = torch.from_numpy(h) if CC else E
            tw = torch.from_numpy(w)
            B = R.bounds([bname])
            key = f"w{k:04d}"
            out[key + "_meta"] = np.array([BOUNDS.index(bname), CC], np.int64)
            out[key + "_pen"] = np.array(p + tuple(vs))
            out[key + "_x"] = x
            out[key + "_g"] = g
            out[key + "_w"] = w
            if CC:
                out[key + "_h"] = h
            out[key + "_reg"] = m.regulariser_grid(torch.from_numpy(x), E, tw, th, *p, vs, B).numpy()
            if p[1]:
                out[key + "_relax"] = m.relax_grid(th, torch.from_numpy(g), torch.from_numpy(x.copy()), tw,
                                                   *p, vs, B, 2).numpy()
            out[key + "_precond"] = m.precond_grid(th, torch.from_numpy(g), E, tw, *p, vs, B).numpy()
            out[key + "_pcg"] = m.pcg_grid(th, torch.from_numpy(g), torch.from_numpy(x.copy() * 0.1), tw,
                                           *p, vs, B, 5, 0.0).numpy()
            k += 1
    f = 0
    shape = (12, 10, 9)
    for bname in ("dft", "dct2"):
        for p in PENALTIES[1:]:
            for use_cg in (False, True):
                g = rng.standard_normal((1, 3) + shape)
                w = rng.uniform(0.2, 2.0, (1, 1) + shape)
                h = rng.uniform(0.5, 2, (1, 6) + shape)
                h[:, 3:] = rng.uniform(-0.2, 0.2, (1, 3) + shape)
                sol = m.fmg_grid(torch.from_numpy(h), torch.from_numpy(g), E, torch.from_numpy(w), *p,
                                 [1., 1., 1.], R.bounds([bname]), 2, 2, 16, use_cg).numpy()
                key = f"f{f:04d}"
                out[key + "_meta"] = np.array([BOUNDS.index(bname), int(use_cg)], np.int64)
                out[key + "_pen"] = np.array(p)
                out[key + "_g"] = g
                out[key + "_h"] = h
                out[key + "_w"] = w
                out[key + "_sol"] = sol
                f += 1
    path = os.path.join(HERE, "solver_rls_golden.npz")
    np.savez_compressed(path, **out)
    print(path, k, f)


if __name__ == "__main__":
    main()
