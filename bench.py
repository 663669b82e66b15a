"""Benchmark of the spline push/pull hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-sweep] [--no-cpu] [--no-solve] [--copy-calibration]

Workload ("pushpull_192"): input 4x8x192^3 float32 U[0,1), grid = identity +
smoothed displacement with 5% of the points out of bounds (seeded,
nitorch_amd/bench_data.py).  One step = one grid_pull and one grid_push,
cubic B-splines (order 3), bound dct2, no extrapolation, inputs resident in
HBM.  `value` = spatial samples processed per second over the step
(2 x 4 x 192^3 samples per step), in Gvox/s.

Also reported: per-op times (and the autograd backward of an order-3 dft
pull on 2x3x128x160x128, both gradients; the cubic dct2 prefilter of the
4x8x192^3 image; its cubic edge-aligned resize to 96^3), the Gauss-Newton
solves of the 160x192x160 system and the SVF exponentiation and its analytic
backward on that lattice, the order x bound sweep of pull/push/count/grad
(orders 0/1/3/5 x dct2/dft/replicate/zero, median of three timed launches
each after a warm-up), the roofline of the dominant kernel (algorithmic bytes per launch /
average launch time, HIP events on the launch stream), and the reference's
own C++ CPU backend (oracle/_ref, built from the reference sources) timed on
a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, published
METRIC = "pull/push Gvox/s + %HBM peak; GN step solve ms to 1e-4 rel residual"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-sweep", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-slices", type=int, default=48,
                   help="CPU baseline sample: first K x-slices of every batch element")
    p.add_argument("--no-solve", action="store_true")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first copy 1 GiB device to device (a known-size read and write for "
                        "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def timed(fn, reps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in ev]
    return sum(ms) / len(ms), ms


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the kernels here run on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is not implemented)"}))
        sys.exit(2)
    import torch
    from nitorch_amd import grid as G
    from nitorch_amd.bench_data import config2_inputs

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    if args.copy_calibration:
        a = torch.ones(256 << 20, device=dev)      # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)                                 # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del a, b
    x_cpu, g_cpu = config2_inputs()
    x = x_cpu.to(dev)
    g = g_cpu.to(dev)
    N, C = x.shape[:2]
    shape = list(x.shape[2:])
    S = g[..., 0].numel()           # spatial samples per op
    inb = ((g >= -0.05) & (g < torch.tensor(shape, device=dev, dtype=g.dtype) - 1 + 0.05)).all(-1)
    oob_frac = 1.0 - inb.float().mean().item()
    gen = torch.Generator().manual_seed(22)
    y_cpu = torch.rand(x_cpu.shape, generator=gen)    # values pushed
    y = y_cpu.to(dev)

    B3, O3 = ["dct2"], [3]
    pull = lambda: G.grid_pull(x, g, B3, O3, 0)
    push = lambda: G.grid_push(y, g, shape, B3, O3, 0)

    # timed region: K steps of (pull + push), bracketed by synchronisations
    for _ in range(args.warmup):
        pull(); push()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pull(); push()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    pull_ms, _ = timed(pull, max(args.steps, 5), 1, stream)
    push_ms, _ = timed(push, max(args.steps, 5), 1, stream)

    # algorithmic bytes per sample (float32): grid 12 + input C*4 + output C*4
    bytes_per_sample = 12 + 8 * C
    gb_launch = bytes_per_sample * S / 1e9
    ops = {
        "pull_o3_dct2": {"ms": pull_ms, "gvox_s": S / pull_ms / 1e6, "hbm_frac": gb_launch / (pull_ms / 1e3) / HBM_PEAK_GBS},
        "push_o3_dct2": {"ms": push_ms, "gvox_s": S / push_ms / 1e6, "hbm_frac": gb_launch / (push_ms / 1e3) / HBM_PEAK_GBS},
    }
    dom = max(ops, key=lambda k: ops[k]["ms"])
    achieved = gb_launch / (ops[dom]["ms"] / 1e3)

    # the same pull / push on deformations far from the identity: the grid
    # recipe with the displacement rescaled to its nominal std (8 voxels RMS),
    # and an affine grid (15 degree rotation about y, zoom 1.1, about the
    # centre); each with the fraction of tiles whose taps do not fit the LDS
    # box (their samples are sorted by first-tap bin and run the binned pass)
    deformed = {}
    for variant in ("rescaled", "affine"):
        if variant == "rescaled":
            gv = config2_inputs(rescale=True)[1].to(dev)
        else:
            gv = affine_grid_192(dev, N)
        pv = lambda: G.grid_pull(x, gv, B3, O3, 0)
        qv = lambda: G.grid_push(y, gv, shape, B3, O3, 0)
        pms, _ = timed(pv, max(args.steps, 5), 1, stream)
        qms, _ = timed(qv, max(args.steps, 5), 1, stream)
        disp = (gv - g_ident(dev, shape)).double()
        disp = disp[((gv >= 0) & (gv <= torch.tensor(shape, device=dev) - 1)).all(-1)]
        for nm, ms, gather in (("pull", pms, True), ("push", qms, False)):
            ops[f"{nm}_o3_dct2_{variant}"] = {"ms": ms, "gvox_s": S / ms / 1e6,
                                              "hbm_frac": gb_launch / (ms / 1e3) / HBM_PEAK_GBS,
                                              "fallback_tiles": fallback_fraction(gv, shape, 3, gather),
                                              "displacement_rms": float(disp.pow(2).mean().sqrt())}
        del gv
    ops["pull_o3_dct2"]["fallback_tiles"] = fallback_fraction(g, shape, 3, True)
    ops["push_o3_dct2"]["fallback_tiles"] = fallback_fraction(g, shape, 3, False)
    ops["pull_o3_dct2"]["displacement_rms"] = float((g - g_ident(dev, shape)).double()[inb].pow(2).mean().sqrt())

    # backward_128: autograd backward of grid_pull (order 3, dft), both gradients
    from nitorch_amd.bench_data import config3_inputs
    x3, g3, d3 = (t.to(dev) for t in config3_inputs())
    x3.requires_grad_(True)
    g3.requires_grad_(True)
    S3 = g3[..., 0].numel()
    bwd = lambda: G.grid_pull_backward(d3, x3, g3, ["dft"], [3], 0)
    bwd_ms, _ = timed(bwd, max(args.steps, 5), 1, stream)
    bps3 = 24 + 12 * x3.shape[1]     # grid + grad_grid 24, input + grad + grad_input 12 per channel
    ops["pull_backward_o3_dft"] = {"ms": bwd_ms, "gvox_s": S3 / bwd_ms / 1e6,
                                   "hbm_frac": bps3 * S3 / (bwd_ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    del x3, g3, d3

    # prefilter_192: interpolating coefficients of the 4x8x192^3 image (cubic,
    # dct2), in place; one HBM read + write of the volume per axis
    xs = x.clone()
    pre = lambda: G.spline_coeff_nd(xs, ["dct2"], [3], 3, True)
    pre_ms, _ = timed(pre, 5, 1, stream)
    ops["prefilter_o3_dct2"] = {"ms": pre_ms, "gvox_s": S / pre_ms / 1e6,
                                "hbm_frac": 3 * 8 * C * S / (pre_ms / 1e3) / 1e9 / HBM_PEAK_GBS}
    del xs

    # resize_192: the 4x8x192^3 image resized by 1/2 (cubic, edge-aligned,
    # dct2) to 96^3; one read of the image and one write of the result
    from nitorch_amd import resize as RZ
    rz = lambda: RZ.resize(x, None, [0.5], ["dct2"], [3], ["edge"])
    rz_ms, _ = timed(rz, 5, 1, stream)
    vout = S // 8
    ops["resize_half_o3_dct2"] = {"ms": rz_ms, "gvox_s": vout / rz_ms / 1e6,
                                  "hbm_frac": 4 * C * (S + vout) / (rz_ms / 1e3) / 1e9 / HBM_PEAK_GBS}

    # pull_96: the reference's CPU-runnable case, 1x1x96^3, order 1, dct2
    from nitorch_amd.bench_data import config1_inputs
    x1c, g1c = config1_inputs()
    x1, g1 = x1c.to(dev), g1c.to(dev)
    S1 = g1[..., 0].numel()
    p1 = lambda: G.grid_pull(x1, g1, ["dct2"], [1], 0)
    p1_ms, _ = timed(p1, max(args.steps, 5), 2, stream)
    ops["pull_96_o1_dct2"] = {"ms": p1_ms, "gvox_s": S1 / p1_ms / 1e6,
                              "hbm_frac": 20 * S1 / (p1_ms / 1e3) / 1e9 / HBM_PEAK_GBS}

    sweep = {}
    if not args.no_sweep:
        for order in (0, 1, 3, 5):
            for bound in ("dct2", "dft", "replicate", "zero"):
                Bb, Oo = [bound], [order]
                fns = {
                    "pull": (lambda: G.grid_pull(x, g, Bb, Oo, 0), 12 + 8 * C),
                    "push": (lambda: G.grid_push(y, g, shape, Bb, Oo, 0), 12 + 8 * C),
                    "count": (lambda: G.grid_count(g, shape, Bb, Oo, 0), 12 + 4),
                    # order 0: the gradient is zero, only the output write is compulsory
                    "grad": (lambda: G.grid_grad(x, g, Bb, Oo, 0), 12 * C if order == 0 else 12 + 4 * C + 12 * C),
                }
                for name, (fn, bps) in fns.items():
                    _, each = timed(fn, 3, 1, stream)
                    ms = sorted(each)[1]
                    sweep[f"{name}_o{order}_{bound}"] = {
                        "ms": round(ms, 4), "gvox_s": round(S / ms / 1e6, 2),
                        "hbm_frac": round(bps * S / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    solve = None
    if not args.no_solve:
        solve = gn_solve(dev)
        solve["svf_exp"] = svf_exp(dev, stream)

    cpu = None
    if not args.no_cpu:
        # GPU outputs on the CPU-baseline sample, for the error statements
        k = args.cpu_slices
        pull_out = pull().cpu()
        ym = torch.zeros_like(y)
        ym[:, :, :k] = y[:, :, :k]
        push_slab = G.grid_push(ym, g, shape, B3, O3, 0).cpu()
        ym.zero_()
        ym[:1, :, :2] = y[:1, :, :2]
        push_o = G.grid_push(ym, g, shape, B3, O3, 0)[:1].cpu()
        del ym
        cpu = cpu_baseline(x_cpu, g_cpu, y_cpu, pull_out, push_slab, push_o, k)
        cpu["pull_96_o1_dct2"] = cpu_pull96(x1c, g1c, p1().cpu())

    line = {
        "metric": METRIC,
        "value": round(2 * S / wall_ms / 1e6, 3),
        "unit": "Gvox/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded U[0,1) image, smoothed random displacement, 5% out of bounds)",
        "config": {"workload": "pushpull_192: grid_pull + grid_push, order 3, dct2, "
                               "input 4x8x192^3, grid 4x192^3x3",
                   "samples_per_op": S, "oob_fraction": round(oob_frac, 4)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dom),
                     "algorithmic_bytes_per_launch": int(bytes_per_sample * S)},
        "ops": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in ops.items()},
        "sweep": sweep,
        "gn_solve": solve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def g_ident(dev, shape):
    import torch
    mesh = torch.meshgrid(*[torch.arange(n, dtype=torch.float32, device=dev) for n in shape], indexing="ij")
    return torch.stack(mesh, -1)


def affine_grid_192(dev, batch, angle_deg=15.0, zoom=1.1, shape=(192, 192, 192)):
    """Voxel coordinates of a rotation by `angle_deg` about the y axis and a
    zoom, about the volume centre: x_in = zoom * R (x_out - c) + c."""
    import math
    import torch
    a = math.radians(angle_deg)
    rot = torch.tensor([[math.cos(a), 0.0, math.sin(a)], [0.0, 1.0, 0.0], [-math.sin(a), 0.0, math.cos(a)]],
                       dtype=torch.float64, device=dev) * zoom
    c = torch.tensor([(n - 1) / 2 for n in shape], dtype=torch.float64, device=dev)
    ijk = g_ident(dev, shape).double() - c
    out = (ijk @ rot.T + c).float()
    return out.expand(batch, *out.shape).contiguous()


def fallback_fraction(grid, shape, order, gather):
    """Fraction of the tiles (4x8x16 samples for gathers, 8x16x16 for
    scatters up to order 3) whose in-bound samples' first taps spread more
    than the LDS box allows (tile extent + 1 per axis); those tiles' samples
    are sorted by first-tap bin and run the binned second pass."""
    import torch
    tile = (4, 8, 16) if (gather or order > 3) else (8, 16, 16)
    N = grid.shape[0]
    tol = 0.05
    lim = torch.tensor(shape, device=grid.device, dtype=grid.dtype) - 1 + tol
    inb = ((grid >= -tol) & (grid < lim)).all(-1)
    if order == 0:
        lo = torch.round(grid)
    elif order == 1:
        lo = torch.floor(grid)
    else:
        lo = torch.floor(grid - (order - 1) * 0.5)
    pad = [(-n) % t for n, t in zip(grid.shape[1:4], tile)]
    big = torch.tensor(2 ** 30, dtype=lo.dtype, device=lo.device)
    fits = torch.ones((), dtype=torch.bool, device=grid.device)
    nt = None
    has = None
    for d in range(3):
        v = lo[..., d]
        mn = torch.where(inb, v, big)
        mx = torch.where(inb, v, -big)
        mn = torch.nn.functional.pad(mn, (0, pad[2], 0, pad[1], 0, pad[0]), value=float(2 ** 30))
        mx = torch.nn.functional.pad(mx, (0, pad[2], 0, pad[1], 0, pad[0]), value=-float(2 ** 30))
        tx, ty, tz = (s // t for s, t in zip(mn.shape[1:], tile))
        mn = mn.reshape(N, tx, tile[0], ty, tile[1], tz, tile[2]).amin((2, 4, 6))
        mx = mx.reshape(N, tx, tile[0], ty, tile[1], tz, tile[2]).amax((2, 4, 6))
        has = mn <= mx
        fits = fits & (mx - mn <= tile[d] + 1)
        nt = has.numel()
    return round(float((has & ~fits).sum()) / nt, 4)


def measured_traffic(op):
    """HBM bytes per launch of the op's kernels, from the committed counter
    runs (profiles/traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE per
    dispatch, divided by what the same counters read for a copy of known
    size); None when no counter run is recorded for this op."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(op)
    return None if rec is None else int(rec["hbm_bytes_per_launch"])


def gn_solve(dev):
    """Gauss-Newton step of the fmg_160 workload (160x192x160x3, bending 0.2,
    membrane 1e-3, absolute 1e-4, lame (0.05, 0.2), dft): plain GPU FMG with
    5 levels (coarsest 10x12x10), 2 V-cycles per level and 4 smoothing steps,
    then V-cycles until the relative residual is <= 1e-4.  Two smoothers:
    multi-colour Gauss-Seidel ('relax') and block-Jacobi PCG ('cg', the
    reference's default).  Times each whole solve after one warm-up."""
    import torch
    from nitorch_amd.bench_data import config4_inputs
    from nitorch_amd import solve as SV
    h, g = config4_inputs()
    h = h.to(dev).movedim(-1, 1).contiguous()
    g = g.to(dev).movedim(-1, 1).contiguous()
    pen = (1e-4, 1e-3, 0.2, 0.05, 0.2)
    out = {"workload": "fmg_160: 1x3x160x192x160, H 6 ch, dft", "tol": 1e-4,
           "levels": 5, "nb_cycles": 2, "nb_iter": 4}
    for name, use_cg in (("relax", False), ("cg", True)):
        kw = dict(tol=1e-4, nb_cycles=2, nb_iter=4, max_levels=5, use_cg=use_cg)
        SV.solve_to_tolerance(h, g, pen, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, rel, ncyc = SV.solve_to_tolerance(h, g, pen, **kw)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        out[name] = {"ms_to_1e-4": round(ms, 3), "rel_residual": rel, "extra_vcycles": ncyc}
    # the same multigrid with grid_pull / grid_push transfers driven from
    # Python (nitorch_amd/fmg.py), one FMG pass: 2 cycles, 4 PCG steps
    from nitorch_amd import fmg as PF
    pf = PF.GridFMG(*pen[:3], pen[3:], bound="dft", nb_cycles=2, nb_iter=4, max_levels=4, optim="cg")
    pf.solve(h, g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = pf.solve(h, g)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    r = g - SV.regulariser_grid_cf(x, h, pen, (1, 1, 1), "dft")
    out["pushpull_fmg"] = {"ms_one_fmg_pass": round(ms, 3), "rel_residual": float(r.norm() / g.norm()),
                           "levels": len(pf.pyrn), "smoother": "cg"}
    best = min(("relax", "cg"), key=lambda k: out[k]["ms_to_1e-4"])
    out["plain_fmg_ms"] = out[best]["ms_to_1e-4"]
    out["plain_fmg_smoother"] = best
    # learned preconditioner (trained weights committed with the package)
    from nitorch_amd import learned as LN
    wpath = LN.default_weights_path()
    if os.path.exists(wpath):
        net = LN.UNet.load(wpath, dev)
        kw = dict(tol=1e-4, max_iter=200)
        LN.pcg_learned(net, h, g, pen, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, rel, it, fb = LN.pcg_learned(net, h, g, pen, **kw)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        out["learned"] = {"ms_to_1e-4": round(ms, 3), "rel_residual": rel, "pcg_iterations": it,
                          "fmg_fallback": fb, "vs_plain_fmg": round(ms / out["plain_fmg_ms"], 3),
                          "method": "flexible PCG, U-Net (3 levels, 16/32/64, BF16 MFMA) warm start + "
                                    "preconditioner on block-Jacobi, FMG fallback"}
    return out


def svf_exp(dev, stream):
    """Scaling and squaring of a smooth 160x192x160x3 velocity (8 steps,
    linear, dft): 8 pulls of the 3-channel field on the HIP kernels."""
    import torch
    from nitorch_amd import svf
    from nitorch_amd.bench_data import smooth
    gen = torch.Generator().manual_seed(5)
    v = smooth(torch.randn((1, 3, 160, 192, 160), generator=gen, dtype=torch.float64), 4.0)
    v = (v * (3.0 / v.abs().max())).movedim(1, -1).float().contiguous().to(dev)
    ms, _ = timed(lambda: svf.exp(v, steps=8), 5, 1, stream)
    g = torch.randn(v.shape, generator=gen).to(dev)
    bms, _ = timed(lambda: svf.exp_backward(v, g, steps=8, rotate_grad=True), 3, 1, stream)
    return {"workload": "1x160x192x160x3 velocity, |v| <= 3 vox, 8 steps, linear, dft", "ms": round(ms, 3),
            "backward_ms": round(bms, 3), "backward": "exp_backward(rotate_grad=True), the anagrad path"}


def host_facts():
    """nproc, CPU model and torch's intra-op thread count of this host."""
    import torch
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "torch_threads": torch.get_num_threads()}


def _maxrel(a, b):
    """max |a - b| / max |b| (norm-relative error)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(float(np.abs(b).max()), 1e-300))


def _elementwise(a, b, bar):
    """max over elements of |a - b| / bar (bar = the push of |values|)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    bar = np.asarray(bar, dtype=np.float64)
    err = np.abs(a - b)
    nz = bar > 0
    if (err[~nz] > 0).any():
        return float("inf")
    return float((err[nz] / bar[nz]).max()) if nz.any() else 0.0


def cpu_baseline(x_cpu, g_cpu, y_cpu, pull_gpu, push_masked, push_o, slices):
    """Reference C++ CPU backend (oracle/_ref, built from the reference
    sources) on a bounded sample of the workload -- every batch element, the
    first `slices` x-slices of the grid, pull and push, order 3, dct2, the
    same input, grid and pushed values as the GPU -- timed as the median of 5
    runs on all host threads, plus one single-thread run of a smaller
    sample.  The GPU results are compared with the reference's on that
    sample (pull: the GPU's full-size output restricted to the slab; push:
    the GPU's push of the values zeroed outside the slab).  Falls back to
    the oracle port (single thread, 4 slices of batch 0) when the reference
    build is absent."""
    import numpy as np
    import torch
    x = x_cpu.contiguous()
    g = g_cpu[:, :slices].contiguous()
    y = y_cpu[:, :, :slices].contiguous()
    S = g[..., 0].numel()
    threads = torch.get_num_threads()
    try:
        from oracle import ref as R
        m = R.load()
    except Exception:
        m = None
    errors = {}
    if m is not None:
        B, I = R.bounds(["dct2"]), R.orders([3])
        tp, tq = [], []
        for _ in range(5):
            t0 = time.perf_counter()
            rp = m.grid_pull(x, g, B, I, 0)
            t1 = time.perf_counter()
            rq = m.grid_push(y, g, list(x.shape[2:]), B, I, 0)
            t2 = time.perf_counter()
            tp.append(t1 - t0)
            tq.append(t2 - t1)
        pull_s, push_s = sorted(tp)[2], sorted(tq)[2]
        errors["pull_vs_reference"] = _maxrel(pull_gpu[:, :, :slices].double().numpy(), rp.double().numpy())
        errors["push_vs_reference"] = _maxrel(push_masked.double().numpy(), rq.double().numpy())
        kind = "reference"
        sample = (f"all {g.shape[0]} batch elements, first {slices} of {g_cpu.shape[1]} x-slices of the grid "
                  f"({S} samples), 8 channels, order 3, dct2, pull + push; median of 5 runs")
        g1 = g[:1, :8].contiguous()
        y1 = y[:1, :, :8].contiguous()
        S1 = g1[..., 0].numel()
        torch.set_num_threads(1)
        try:
            u0 = time.perf_counter()
            m.grid_pull(x[:1], g1, B, I, 0)
            m.grid_push(y1, g1, list(x.shape[2:]), B, I, 0)
            u1 = time.perf_counter()
        finally:
            torch.set_num_threads(threads)
        single = {"value": round(2 * S1 / (u1 - u0) / 1e9, 6), "unit": "Gvox/s", "cores": 1,
                  "sample": f"batch 0, first 8 x-slices ({S1} samples), pull + push, one run"}
    else:
        from oracle import pushpull as O
        x, g, y = x[:1], g[:1, :4], y[:1, :, :4]
        S = g[..., 0].numel()
        t0 = time.perf_counter()
        O.grid_pull(x.numpy(), g.numpy(), "dct2", 3, 0)
        t1 = time.perf_counter()
        O.grid_push(y.numpy(), g.numpy(), x.shape[2:], "dct2", 3, 0)
        t2 = time.perf_counter()
        pull_s, push_s = t1 - t0, t2 - t1
        kind, threads = "port", 1
        sample = f"batch 0, first 4 x-slices ({S} samples), single-threaded oracle port, one run"
        single = None
    # error against the FP64 oracle: batch 0, first 2 x-slices
    from oracle import pushpull as O
    k = min(2, slices)
    x64 = x_cpu[:1].double().numpy()
    g64 = g_cpu[:1, :k].double().numpy()
    y64 = y_cpu[:1, :, :k].double().numpy()
    errors["pull_vs_fp64_oracle"] = _maxrel(pull_gpu[:1, :, :k].double().numpy(), O.grid_pull(x64, g64, "dct2", 3, 0))
    shp = tuple(x_cpu.shape[2:])
    errors["push_vs_fp64_oracle_elementwise"] = _elementwise(push_o.double().numpy(),
                                                             O.grid_push(y64, g64, shp, "dct2", 3, 0),
                                                             O.grid_push(abs(y64), g64, shp, "dct2", 3, 0))
    errors["definition"] = ("pull: max|gpu - ref| / max|ref| over the sample; push vs reference: the same "
                            "norm-relative error; push vs FP64 oracle: max over elements of |gpu - ref| / "
                            "push(|values|) (batch 0, first 2 x-slices)")
    return {"value": round(2 * S / (pull_s + push_s) / 1e9, 5), "unit": "Gvox/s", "cores": threads,
            "kind": kind, "sample": sample, "host": host_facts(),
            "pull_s": round(pull_s, 3), "push_s": round(push_s, 3), "single_thread": single,
            "errors": errors}


def cpu_pull96(x, g, gpu_out):
    """The reference CPU backend on the whole pull_96 case (median of 5),
    and the GPU output's error against it."""
    import torch
    try:
        from oracle import ref as R
        m = R.load()
    except Exception:
        m = None
    if m is None:
        return None
    B, I = R.bounds(["dct2"]), R.orders([1])
    ref = m.grid_pull(x, g, B, I, 0)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        m.grid_pull(x, g, B, I, 0)
        ts.append(time.perf_counter() - t0)
    ms = sorted(ts)[2] * 1e3
    return {"ms": round(ms, 3), "gvox_s": round(g[..., 0].numel() / ms / 1e6, 4), "cores": torch.get_num_threads(),
            "kind": "reference", "error_vs_reference": _maxrel(gpu_out.double().numpy(), ref.double().numpy())}


if __name__ == "__main__":
    main()
