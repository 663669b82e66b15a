"""GPU parity of the HIP solver kernels against the solver oracle
(intended semantics: every colour class visited, per-channel FMG rescaling).

Tolerances (norm-relative, max|a-b| / max|b|): FP64 1e-10 (relax / FMG
1e-9: several sweeps), FP32 1e-5 for single operators, 1e-4 for FMG
solutions.  The GPU relax visits colour classes in the oracle's order and
updates each class in parallel (independent sets), so Gauss-Seidel iterates
match to rounding.
"""
import numpy as np
import pytest
import torch

from oracle import solver as S

pytestmark = pytest.mark.gpu
DEV = "cuda"

PENS = [
    dict(absolute=1e-3),
    dict(absolute=1e-4, membrane=0.1),
    dict(absolute=1e-4, membrane=1e-3, bending=0.2),
    dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2)),
    dict(membrane=0.1, lame=(0.3, 0.1)),
]


def pen_vec(p):
    lame = p.get("lame", (0, 0))
    return (p.get("absolute", 0), p.get("membrane", 0), p.get("bending", 0), lame[0], lame[1])


def rel(a, b):
    a = a.detach().cpu().double().numpy() if torch.is_tensor(a) else a
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def system(rng, shape, CC=6, N=1):
    g = rng.standard_normal((N, 3) + shape)
    h = None
    if CC:
        h = rng.uniform(0.5, 2, (N, CC) + shape)
        if CC == 6:
            h[:, 3:] = rng.uniform(-0.2, 0.2, (N, 3) + shape)
    return g, h


def T(a, dt):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dt)


@pytest.fixture(scope="module")
def sv():
    from nitorch_amd import solve
    return solve


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
@pytest.mark.parametrize("bound", ["dft", "dct2", "replicate", "dst2", "zero", "dct1"])
def test_regulariser_vs_oracle(sv, dt, bound):
    rng = np.random.default_rng(1)
    tol = 1e-5 if dt == torch.float32 else 1e-10
    for shape in [(7, 6, 9), (12, 10, 11)]:
        for p in PENS:
            for CC in (0, 1, 3, 6):
                for N in (1, 2):
                    x, h = system(rng, shape, CC, N)
                    vs = (1.0, 1.3, 0.8)
                    out = sv.regulariser_grid_cf(T(x, dt), T(h, dt), pen_vec(p), vs, bound)
                    ref = S.regulariser_grid(x, h, voxel_size=vs, bound=bound, **p)
                    assert rel(out, ref) <= tol, (shape, p, CC, N, bound)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_relax_vs_oracle(sv, dt):
    rng = np.random.default_rng(2)
    # single-block and multi-launch paths; sizes with and without remainder
    # colour classes on circulant axes
    for shape in [(9, 8, 10), (12, 6, 18), (40, 36, 34), (30, 24, 36)]:
        for p in PENS:
            for bound in ("dft", "dct2"):
                g, h = system(rng, shape)
                x0 = rng.standard_normal(g.shape) * 0.1
                x = T(x0, dt)
                sv.relax_grid_cf(T(g, dt), x, T(h, dt), pen_vec(p), (1, 1, 1), bound, nb_iter=2)
                ref = S.relax_grid(g, h, x0, bound=bound, nb_iter=2, product_order=True, **p)
                tol = 2e-5 if dt == torch.float32 else 1e-9
                assert rel(x, ref) <= tol, (shape, p, bound)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_transfers_vs_oracle(sv, dt):
    rng = np.random.default_rng(3)
    tol = 1e-6 if dt == torch.float32 else 1e-12
    for shape in [(7, 6, 5), (16, 9, 12), (3, 3, 3), (5, 1, 4)]:
        for bound in ("dct2", "dft", "replicate", "zero", "dst2", "dct1"):
            x = rng.standard_normal((2, 3) + shape)
            cs = tuple(1 + (s - 1) // 2 for s in shape)
            out = sv.fmg_restriction(T(x, dt), torch.empty((2, 3) + cs, dtype=dt, device=DEV), bound)
            assert rel(out, S.restrict(x, cs, bound)) <= tol
            xc = rng.standard_normal((2, 3) + cs)
            out = sv.fmg_prolongation(T(xc, dt), torch.empty((2, 3) + shape, dtype=dt, device=DEV), bound)
            assert rel(out, S.prolong(xc, shape, bound)) <= tol


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_transfers_tiled_vs_oracle(sv, dt):
    """Transfers of at least 2^16 outputs run the LDS-tiled kernel (odd and
    even sizes, folded and signed bounds)."""
    rng = np.random.default_rng(13)
    tol = 1e-6 if dt == torch.float32 else 1e-12
    for shape in [(64, 50, 45), (48, 64, 33)]:
        for bound in ("dct2", "dft", "dst2", "zero"):
            x = rng.standard_normal((2, 3) + shape)
            cs = tuple(1 + (s - 1) // 2 for s in shape)
            out = sv.fmg_restriction(T(x, dt), torch.empty((2, 3) + cs, dtype=dt, device=DEV), bound)
            assert rel(out, S.restrict(x, cs, bound)) <= tol, ("restrict", shape, bound)
            xc = rng.standard_normal((2, 3) + cs)
            out = sv.fmg_prolongation(T(xc, dt), torch.empty((2, 3) + shape, dtype=dt, device=DEV), bound)
            assert rel(out, S.prolong(xc, shape, bound)) <= tol, ("prolong", shape, bound)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_fmg_vs_oracle(sv, dt):
    rng = np.random.default_rng(4)
    for shape in [(12, 10, 9), (24, 20, 18)]:
        for p in PENS[2:]:
            for bound in ("dft", "dct2"):
                g, h = system(rng, shape)
                x = sv.fmg_grid_cf(T(h, dt), T(g, dt), None, pen_vec(p), (1, 1, 1), bound, 2, 2)
                ref = S.fmg_grid(h, g, bound=bound, nb_cycles=2, nb_iter=2, product_order=True, **p)
                tol = 1e-4 if dt == torch.float32 else 1e-9
                assert rel(x, ref) <= tol, (shape, p, bound)


@pytest.mark.parametrize("case", [
    ((64, 64, 64), 1, "dft", 3, torch.float32),      # 8 classes, K = 8 slots a row
    ((64, 64, 64), 1, "dft", 3, torch.float64),
    ((40, 48, 40), 4, "dct2", 2, torch.float32),     # 7 classes, partial last slot, batch of 4
    ((64, 64, 64), 1, "dft", 1, torch.float32),      # membrane only: red/black
])
def test_fmg_class_major_vs_oracle(sv, case):
    """Levels of at least 2^16 voxels (N V) relax in the class-major layout
    (here the finest level of each case): the FMG iterate equals the
    oracle's (same colour order)."""
    shape, N, bound, ip, dt = case
    rng = np.random.default_rng(40 + ip)
    p = PENS[ip]
    g, h = system(rng, shape, N=N)
    x = sv.fmg_grid_cf(T(h, dt), T(g, dt), None, pen_vec(p), (1, 1, 1), bound, 2, 2)
    ref = S.fmg_grid(h, g, bound=bound, nb_cycles=2, nb_iter=2, product_order=True, **p)
    tol = 1e-4 if dt == torch.float32 else 1e-9
    assert rel(x, ref) <= tol, (shape, N, bound, p)


def test_relax_deterministic_with_wrap(sv):
    """Circulant axes whose length is not a multiple of the colour period:
    repeated relaxations give bit-identical results (no races)."""
    rng = np.random.default_rng(7)
    g, h = system(rng, (41, 37, 40))
    outs = []
    for _ in range(2):
        x = torch.zeros((1, 3, 41, 37, 40), dtype=torch.float32, device=DEV)
        sv.relax_grid_cf(T(g, torch.float32), x, T(h, torch.float32), pen_vec(PENS[3]), (1, 1, 1), "dft", 3)
        outs.append(x.cpu())
    assert torch.equal(outs[0], outs[1])


def test_solve_to_tolerance(sv):
    rng = np.random.default_rng(5)
    shape = (40, 48, 40)
    g, h = system(rng, shape)
    p = PENS[3]
    gt, ht = T(g, torch.float32), T(h, torch.float32)
    x, r, ncyc = sv.solve_to_tolerance(ht, gt, pen_vec(p), tol=1e-4, nb_cycles=2, nb_iter=4)
    assert r <= 1e-4
    # independent check of the residual in FP64 with the oracle operator
    res = g - S.regulariser_grid(x.cpu().double().numpy(), h, **p)
    assert np.linalg.norm(res) / np.linalg.norm(g) <= 2e-4


def test_channel_last_api(sv):
    rng = np.random.default_rng(6)
    shape = (10, 12, 9)
    g, h = system(rng, shape)
    p = dict(absolute=1e-4, membrane=1e-3, bending=0.2, lame=(0.05, 0.2))
    gl = T(np.moveaxis(g, 1, -1), torch.float64)           # (N, X, Y, Z, 3)
    hl = T(np.moveaxis(h, 1, -1), torch.float64)
    out = sv.regulariser_grid(gl, hessian=hl, **p)
    assert out.shape == gl.shape
    ref = np.moveaxis(S.regulariser_grid(g, h, **p), 1, -1)
    assert rel(out, ref) < 1e-10
    sol = sv.fmg_grid(hl, gl, **p, nb_cycles=2, nb_iter=2, solver="relax")
    ref = np.moveaxis(S.fmg_grid(h, g, nb_cycles=2, nb_iter=2, product_order=True, **p), 1, -1)
    assert rel(sol, ref) < 1e-9
    sol = sv.fmg_grid(hl, gl, **p, nb_cycles=2, nb_iter=2)          # reference default: PCG smoother
    ref = np.moveaxis(S.fmg_grid(h, g, nb_cycles=2, nb_iter=2, use_cg=True, **p), 1, -1)
    assert rel(sol, ref) < 1e-9
    sol = sv.pcg_grid(gl, hessian=hl, **p, nb_iter=6, tol=0)
    ref = np.moveaxis(S.pcg_grid(h, g, nb_iter=6, **p)[0], 1, -1)
    assert rel(sol, ref) < 1e-9
    sol = sv.precond_grid(gl, hessian=hl, **p)
    ref = np.moveaxis(S.precond_grid(g, h, **p), 1, -1)
    assert rel(sol, ref) < 1e-12
    # solve_grid_fmg without spatial penalties: closed form
    sol = sv.solve_grid_fmg(hl, gl, absolute=0.5)
    Hm = np.zeros(h.shape[:1] + h.shape[2:] + (3, 3))
    hh = np.moveaxis(h, 1, -1)
    for i, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        Hm[..., a, b] = hh[..., i]
        Hm[..., b, a] = hh[..., i]
    Hm += 0.5 * np.eye(3)
    ref = np.linalg.solve(Hm, np.moveaxis(g, 1, -1)[..., None])[..., 0]
    assert rel(sol, ref) < 1e-10


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_precond_vs_oracle(sv, dt):
    rng = np.random.default_rng(8)
    tol = 1e-6 if dt == torch.float32 else 1e-12
    for shape in [(7, 6, 9), (16, 17, 5)]:
        for p in PENS:
            for CC in (0, 1, 3, 6):
                for N in (1, 2):
                    g, h = system(rng, shape, CC, N)
                    vs = (1.0, 1.3, 0.8)
                    out = sv.precond_grid_cf(T(h, dt), T(g, dt), pen_vec(p), vs)
                    ref = S.precond_grid(g, h, voxel_size=vs, **p)
                    assert rel(out, ref) <= tol, (shape, p, CC, N)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_pcg_vs_oracle(sv, dt):
    """Fixed-iteration PCG (tol 0, the FMG smoother) for all Hessian layouts,
    bounds and batch sizes; sizes below and above one work-group per batch."""
    rng = np.random.default_rng(9)
    for shape in [(7, 6, 5), (20, 18, 22)]:
        for p in PENS[1:]:
            for bound in ("dft", "dct2", "zero"):
                for CC, N in ((6, 1), (3, 2), (1, 1), (0, 2)):
                    g, h = system(rng, shape, CC, N)
                    if CC == 0 and bound == "dft" and p.get("absolute", 0) == 0:
                        continue    # singular system
                    x0 = rng.standard_normal(g.shape) * 0.1
                    x = T(x0, dt)
                    _, it = sv.pcg_grid_cf(T(h, dt), T(g, dt), x, pen_vec(p), (1, 1.2, 0.9), bound, nb_iter=4)
                    assert it == 4
                    ref, _ = S.pcg_grid(h, g, x0, voxel_size=(1, 1.2, 0.9), bound=bound, nb_iter=4, **p)
                    tol = 1e-4 if dt == torch.float32 else 1e-9
                    assert rel(x, ref) <= tol, (shape, p, bound, CC, N)


def test_pcg_large_vs_oracle(sv):
    """Fields of at least 2^22 elements (the LDS-tiled PCG start and
    operator, the chunked direction update; batch of 2, so a chunk can span
    two batch elements), both dtypes against one oracle solve."""
    rng = np.random.default_rng(19)
    shape = (64, 96, 120)
    p = PENS[3]
    g, h = system(rng, shape, 6, 2)
    x0 = rng.standard_normal(g.shape) * 0.1
    ref, _ = S.pcg_grid(h, g, x0, voxel_size=(1, 1.2, 0.9), bound="dft", nb_iter=4, **p)
    for dt, tol in ((torch.float32, 1e-4), (torch.float64, 1e-9)):
        x = T(x0, dt)
        _, it = sv.pcg_grid_cf(T(h, dt), T(g, dt), x, pen_vec(p), (1, 1.2, 0.9), "dft", nb_iter=4)
        assert it == 4
        assert rel(x, ref) <= tol, dt


def test_pcg_early_stop_vs_oracle(sv):
    """tol > 0: same stopping iteration and iterate as the oracle (FP64)."""
    rng = np.random.default_rng(10)
    for shape in [(9, 8, 10), (24, 20, 16)]:
        for p in PENS[1:4]:
            for bound in ("dct2", "replicate"):
                g, h = system(rng, shape, 6, 1)
                x, it = sv.pcg_grid_cf(T(h, torch.float64), T(g, torch.float64), None, pen_vec(p), (1, 1, 1),
                                       bound, nb_iter=60, tol=1e-3)
                ref, it_ref = S.pcg_grid(h, g, bound=bound, nb_iter=60, tol=1e-3, **p)
                assert it == it_ref, (shape, p, bound)
                assert rel(x, ref) <= 1e-9, (shape, p, bound)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_fmg_cg_vs_oracle(sv, dt):
    rng = np.random.default_rng(11)
    for shape in [(12, 10, 9), (24, 20, 18)]:
        for p in PENS[2:]:
            for bound in ("dft", "dct2"):
                g, h = system(rng, shape)
                x = sv.fmg_grid_cf(T(h, dt), T(g, dt), None, pen_vec(p), (1, 1, 1), bound, 2, 2, use_cg=True)
                ref = S.fmg_grid(h, g, bound=bound, nb_cycles=2, nb_iter=2, use_cg=True, **p)
                tol = 2e-4 if dt == torch.float32 else 1e-9
                assert rel(x, ref) <= tol, (shape, p, bound)
