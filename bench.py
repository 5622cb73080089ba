#!/usr/bin/env python3
"""Throughput of one spread + one interpolate of Lagrangian points on one MI355X.

Workload (default): 3-D periodic unit cube, 512^3 cells, side-centred (MAC)
grid, IB_6 kernel, 8,388,608 points: half uniform in the domain, half in a
spherical shell of radius 0.3 and thickness 2h about the centre.  Grid and
force/velocity data FP32 with compensated sums (positions and stencil
selection FP64).  One step = LDataManager::spread into a zeroed field (one
launch) + LDataManager::interp from a ghost-filled field (one launch), points
binned beforehand (binning belongs to regrid in the reference; its time is
reported separately).  The plain FP64 form runs the same workload for the
FP32/FP64 ratio and for the error of the FP32 form.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]

Prints one JSON line.  The CPU baseline is the C oracle (the reference's FP64
loops restated) timed on a bounded sample of the same points on one host core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Lagrangian points spread+interpolated/s (FP32-batched vs FP64) @<=1e-6 err; HBM GB/s"


TILE = None  # --tile


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--points", type=int, default=1 << 23)
    ap.add_argument("--kernel", default="IB_6")
    ap.add_argument("--cpu-sample", type=int, default=400000, help="points of the one-core CPU baseline")
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="threads of the all-cores CPU baseline (capped by the host's cores; 0: skip it)")
    ap.add_argument("--cpu-sample-cores", type=int, default=8388608,
                    help="points of the all-cores CPU baseline (at most the workload's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="default workload only: skip the AMR-level and multi-structure summaries")
    ap.add_argument("--point-order", choices=["regrid", "generated"], default="regrid",
                    help="regrid: the points in LDataManager's order after a redistribution (the reference's "
                         "steady state); generated: the generator's order, random in space")
    ap.add_argument("--workload", choices=["shell512", "amr", "multi"], default="shell512",
                    help="shell512: periodic 512^3 IB_6 (default); amr: multi-patch fine level, BSPLINE_3; "
                         "multi: 64 shells on 384^3 (IB_4 and BSPLINE_3), 100 graph-captured steps")
    ap.add_argument("--tile", default=None,
                    help="X,Y,Z: FP32C spread tile shape instead of the autotuner's choice (the autotuner still "
                         "runs and reports its times)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 2 GiB (calibrates the HBM byte counters)")
    return ap.parse_args()


def copy_probe(torch):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of this
    dispatch divided by 2 GiB is the counter scale factor for the traffic figures."""
    nbytes = 2 << 30
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "dispatches": 3}), flush=True)


def counters_from_profile(kernel, workload_tag):
    """VALU-busy and LDS bank-conflict fractions of a kernel from the committed
    rocprofv3 counter summary of this workload ("Derived" table: the spread
    kernel's column, then the interpolation's), or None."""
    path = os.path.join(ROOT, "profiles", "counters_%s.md" % workload_tag)
    if not os.path.exists(path):
        return None
    col = 0 if kernel.startswith("k_spread") else 1
    out = {}
    with open(path) as fh:
        for line in fh:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 3 and cells[0] == "counter":
                # the summary's kernels (`name<template args>`) must be the ones measured now
                if cells[1 + col].strip("`").split("<")[0] != kernel:
                    return None
            elif len(cells) == 3 and cells[0].startswith("VALU busy"):
                out["valu_busy_frac"] = float(cells[1 + col].rstrip(" %")) / 100.0
            elif len(cells) == 3 and cells[0].startswith("LDS bank-conflict rate"):
                out["lds_bank_conflict_frac"] = float(cells[1 + col].rstrip(" %")) / 100.0
    return out or None


def traffic_from_profile(workload_tag):
    """HBM bytes per dispatch of the dominant kernel, from the committed
    rocprofv3 --pmc summary of this workload (FETCH_SIZE and WRITE_SIZE, each
    scaled by the same counter on a plain copy of known size).  None if absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.txt" % workload_tag)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("TRAFFIC_JSON "):
                return json.loads(line[len("TRAFFIC_JSON "):])
    return None


def make_points(torch, n, N, seed):
    """Half uniform in [0,1)^3, half in the shell |x - c| in [0.3 - h, 0.3 + h)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    h = 1.0 / N
    nu = n // 2
    ns = n - nu
    Xu = torch.rand((nu, 3), generator=g, device="cuda", dtype=torch.float64)
    d = torch.randn((ns, 3), generator=g, device="cuda", dtype=torch.float64)
    d = d / d.norm(dim=1, keepdim=True)
    r = 0.3 - h + 2 * h * torch.rand((ns, 1), generator=g, device="cuda", dtype=torch.float64)
    Xs = 0.5 + r * d
    X = torch.cat([Xu, Xs]).contiguous()
    X = torch.remainder(X, 1.0)
    F = torch.randn((n, 3), generator=g, device="cuda", dtype=torch.float64)
    return X, F


def spread_kernel_name(shape):
    """The 3-D spread kernel of a tile shape: z extents of 48 or more are swept
    z segments (FP32C), shorter ones fixed register columns."""
    return "k_spread3s" if shape[2] >= 48 else "k_spread3"


def touched_faces(torch, X, N, kernel, periodic=True):
    """T: distinct (axis, face) grid values in the union of all point stencils
    (periodic DOFs, or the faces of the N^3 index space and its border when not
    periodic), counted exactly on the GPU."""
    W = {"IB_4": 4, "IB_6": 6, "BSPLINE_3": 3}[kernel]
    h = 1.0 / N
    total = 0
    for a in range(3):
        starts = []
        for d in range(3):
            xl = -0.5 * h if d == a else 0.0
            t = (X[:, d] - xl) / h
            if kernel == "BSPLINE_3":
                s = torch.floor(t) - 1
            else:
                s = torch.floor(t + 0.5) - (W // 2)   # NINT for t >= 0
            starts.append(s.to(torch.int64))
        # not periodic: indices shifted by W into an (N + 2W)^3 box, no wrap
        M, off = (N, 0) if periodic else (N + 2 * W, W)
        wrap = (lambda v: torch.remainder(v, N)) if periodic else (lambda v: v + off)
        mask = torch.zeros(M * M * M, dtype=torch.bool, device="cuda")
        for k2 in range(W):
            i2 = wrap(starts[2] + k2)
            for k1 in range(W):
                i1 = wrap(starts[1] + k1)
                base = (i2 * M + i1) * M
                for k0 in range(W):
                    mask[base + wrap(starts[0] + k0)] = True
        total += int(mask.sum().item())
        del mask
    return total


def time_redistribute(torch, ia, X, F, N, kernel, reps=3):
    """ibamr_plan_redistribute on the points (a regrid: wrap into the periodic
    box, owner patch, new order) plus the permutation of X (float64) and F
    (float32) into the new order; mean ms of reps runs (each ends in a sync)."""
    level = ia.Level.uniform_periodic(N, 3)
    n = X.shape[0]
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP32C, max_points=n)
    Xw = X.clone()
    F32 = F.to(torch.float32).contiguous()
    Xp, Fp = torch.empty_like(Xw), torch.empty_like(F32)
    lag = torch.arange(n, dtype=torch.int32, device="cuda").flip(0).contiguous()
    plan.redistribute(Xw, lag)
    torch.cuda.synchronize()
    import time
    t0 = time.perf_counter()
    for _ in range(reps):
        order, _, _ = plan.redistribute(Xw, lag)
        ia.permute_rows(Xp, Xw, order)
        ia.permute_rows(Fp, F32, order)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    plan.close()
    return ms


def regrid_order(torch, ia, X, F, N, kernel):
    """The points in the order LDataManager keeps its local Lagrangian data after
    a regrid (LDataManager::endDataRedistribution: patch, cell in box order,
    Lagrangian index), computed by ibamr_plan_redistribute; the generator's
    order is random in space."""
    level = ia.Level.uniform_periodic(N, 3)
    n = X.shape[0]
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=n)
    Xw = X.clone()
    order, _, _ = plan.redistribute(Xw, torch.arange(n, dtype=torch.int32, device="cuda"))
    o = order.long()
    torch.cuda.synchronize()
    plan.close()
    return Xw[o].contiguous(), F[o].contiguous()


def run_form(torch, ia, precision, X, F, N, kernel, steps, warmup):
    dt = torch.float64 if precision == ia.FP64 else torch.float32
    level = ia.Level.uniform_periodic(N, 3)
    n = X.shape[0]
    plan = ia.LevelPlan(level, kernel, ia.SIDE, precision, max_points=n)
    g = plan.data_gcw
    f = [torch.zeros(ia.array_shape(level, 0, g, ia.SIDE, axis=a), dtype=dt, device="cuda") for a in range(3)]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234)
    u = [(2 * torch.rand(ia.array_shape(level, 0, g, ia.SIDE, axis=a), generator=gen, device="cuda",
                         dtype=torch.float64) - 1).to(dt) for a in range(3)]
    plan.bind(0, f)
    plan.bind(1, u)
    s = torch.cuda.current_stream()
    plan.fill_periodic_ghosts(1, stream=s)
    Fd = F.to(dt).contiguous()
    U = torch.zeros((n, 3), dtype=dt, device="cuda")
    eb0, eb1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    eb0.record(s)
    plan.bin(X, stream=s, check=False)
    eb1.record(s)
    plan.check(stream=s)
    torch.cuda.synchronize()
    bin_ms = eb0.elapsed_time(eb1)
    # spread tile height chosen by the library's autotuner on these points (untimed)
    tune = plan.autotune(0, Fd, X, reps=3, stream=s)
    if TILE and precision != ia.FP64:
        plan.tile_shape = TILE

    def step(ev=None):
        if ev:
            ev[0].record(s)
        plan.spread(0, Fd, X, mode=ia.OVERWRITE, stream=s)
        if ev:
            ev[1].record(s)
        plan.interp(1, U, None, stream=s)   # same positions: reuse the spread's bin-ordered copy
        if ev:
            ev[2].record(s)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    plan.set_timing(True)   # HIP events on stream s right around the two main kernels
    t0 = time.perf_counter()
    for k in range(steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = plan.kernel_times()
    plan.set_timing(False)
    sp = [e[0].elapsed_time(e[1]) for e in evs]
    it = [e[1].elapsed_time(e[2]) for e in evs]
    return {"plan": plan, "tune": tune, "f": f, "u": u, "U": U, "wall_s": wall, "spread_ms": sum(sp) / len(sp),
            "interp_ms": sum(it) / len(it), "bin_ms": bin_ms, "info": plan.info(),
            "spread_kernel_ms": kt["spread_ms"], "interp_kernel_ms": kt["interp_ms"]}


def host_oracle():
    """Builds the oracle for this host (oracle/Makefile `native`: -O3
    -march=native -ffp-contract=off) into a temporary directory and points the
    oracle loader at it; returns a description of the build (flags, CPU
    model).  A build that fails leaves the plain -O2 library in use and says
    so."""
    import subprocess
    import tempfile

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    odir = os.path.join(ROOT, "oracle")
    out = os.path.join(tempfile.mkdtemp(prefix="ibamr_oracle_"), "liboracle.so")
    try:
        subprocess.run(["make", "-s", "-C", odir, "native", "NATIVE_OUT=" + out], check=True, capture_output=True,
                       timeout=300)
        flags = subprocess.run(["make", "-s", "-C", odir, "native-flags"], check=True, capture_output=True,
                               text=True, timeout=60).stdout.strip()
        os.environ["IBAMR_ORACLE_LIB"] = out
    except (subprocess.SubprocessError, OSError) as e:
        flags = "-O2 -ffp-contract=off (host build failed: %s)" % type(e).__name__
    return {"cpu_model": model, "build": "gcc " + flags}


def cpu_baseline_cores(X_np, F_np, N, kernel, sample, threads, seed=6):
    """The oracle's FP64 spread + interp on `threads` host cores: the periodic
    grid cut into z slabs, one patch per thread (the reference's one patch per
    MPI rank), each thread spreading the points of its patch's Lagrangian ghost
    box (with periodic images) into its ghosted slab and interpolating its
    interior points, all slabs at once (the C oracle releases the GIL)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rng = np.random.default_rng(seed)
    sel = rng.choice(X_np.shape[0], size=min(sample, X_np.shape[0]), replace=False)
    X = np.ascontiguousarray(X_np[sel])
    F = np.ascontiguousarray(F_np[sel])
    h = 1.0 / N
    g = O.min_ghost_width(kernel)
    cell = O.get_cell_index(X, [0.0] * 3, [1.0] * 3, [h] * 3, [0] * 3, [N - 1] * 3)
    slabs = []
    for k in range(threads):
        z0, z1 = (k * N) // threads, ((k + 1) * N) // threads
        il, iu = [0, 0, z0], [N - 1, N - 1, z1 - 1]
        f = [np.zeros(O.side_axis_shape(il, iu, [g] * 3, a)) for a in range(3)]
        idx, sh = O.build_patch_entries(cell, il, iu, g, [N] * 3, [0] * 3, [h] * 3, [8] * 3)
        iidx, ish = O.build_patch_entries(cell, il, iu, 0, [N] * 3, [0] * 3, [h] * 3, [8] * 3, interior_only=True)
        slabs.append((il, iu, [0.0, 0.0, z0 * h], f, idx, sh, iidx, ish))

    def work(sl):
        il, iu, xlo, f, idx, sh, iidx, ish = sl
        O.spread_side(kernel, [h] * 3, xlo, il, iu, [g] * 3, f, X, F, indices=idx, Xshift=sh)
        O.interp_side(kernel, [h] * 3, xlo, il, iu, [g] * 3, f, X, indices=iidx, Xshift=ish)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(work, slabs))
        t = time.perf_counter() - t0
    return {"value": X.shape[0] / t, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": "%d of the %d workload points (%s), full %d^3 side grid cut into %d z-slab "
                      "patches, one thread each, FP64 spread+interp, %.2f s"
                      % (X.shape[0], X_np.shape[0],
                         "all" if X.shape[0] == X_np.shape[0] else "uniform random subset", N, threads, t)}


def cpu_baseline(X_np, F_np, N, kernel, sample, seed=5):
    """The oracle's FP64 spread + interp (the reference CPU loops, one core) on a
    random sample of the workload's points, on the full grid."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rng = np.random.default_rng(seed)
    sel = rng.choice(X_np.shape[0], size=min(sample, X_np.shape[0]), replace=False)
    X = np.ascontiguousarray(X_np[sel])
    F = np.ascontiguousarray(F_np[sel])
    h = 1.0 / N
    g = O.min_ghost_width(kernel)
    il, iu = [0] * 3, [N - 1] * 3
    f = [np.empty(O.side_axis_shape(il, iu, [g] * 3, a)) for a in range(3)]
    for a in f:
        a.fill(0.0)
    cell = O.get_cell_index(X, [0.0] * 3, [1.0] * 3, [h] * 3, il, iu)
    idx, sh = O.build_patch_entries(cell, il, iu, g, [N] * 3, il, [h] * 3, [8] * 3)
    iidx, ish = O.build_patch_entries(cell, il, iu, 0, [N] * 3, il, [h] * 3, [8] * 3, interior_only=True)
    t0 = time.perf_counter()
    O.spread_side(kernel, [h] * 3, [0.0] * 3, il, iu, [g] * 3, f, X, F, indices=idx, Xshift=sh)
    O.interp_side(kernel, [h] * 3, [0.0] * 3, il, iu, [g] * 3, f, X, indices=iidx, Xshift=ish)
    t = time.perf_counter() - t0
    return {"value": X.shape[0] / t, "unit": "points/s", "cores": 1, "kind": "port",
            "sample": "%d of the %d workload points (uniform random subset), full %d^3 side grid, FP64 spread+interp, %.1f s"
                      % (X.shape[0], X_np.shape[0], N, t)}


def run_amr_form(torch, ia, precision, level, X, F, kernel, steps, warmup, seed=77):
    """Spread (overwrite) + interp on every patch of a multi-patch level, one launch each."""
    import numpy as np

    dt = torch.float64 if precision == ia.FP64 else torch.float32
    n = X.shape[0]
    plan = ia.LevelPlan(level, kernel, ia.SIDE, precision, max_points=n)
    g = plan.data_gcw
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    f, u = [], []
    for p in range(len(level.patches)):
        for a in range(3):
            shp = ia.array_shape(level, p, g, ia.SIDE, axis=a)
            f.append(torch.zeros(shp, dtype=dt, device="cuda"))
            u.append((2 * torch.rand(shp, generator=gen, device="cuda", dtype=torch.float64) - 1).to(dt))
    plan.bind(0, f)
    plan.bind(1, u)
    s = torch.cuda.current_stream()
    Fd = F.to(dt).contiguous()
    U = torch.zeros((n, 3), dtype=dt, device="cuda")
    plan.bin(X, stream=s)
    tune = plan.autotune(0, Fd, X, reps=3, stream=s)  # spread tile height for these patches and points
    if TILE and precision != ia.FP64:
        plan.tile_shape = TILE
    for _ in range(warmup):
        plan.spread(0, Fd, X, mode=ia.OVERWRITE, stream=s)
        plan.interp(1, U, X, stream=s)
    torch.cuda.synchronize()
    plan.set_timing(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        plan.spread(0, Fd, X, mode=ia.OVERWRITE, stream=s)
        plan.interp(1, U, X, stream=s)
    e1.record(s)
    torch.cuda.synchronize()
    kt = plan.kernel_times()
    plan.set_timing(False)
    return {"f": f, "U": U, "ms": e0.elapsed_time(e1) / steps, "kt": kt, "info": plan.info(), "tune": tune,
            "shape": plan.tile_shape}


def bench_amr(args, torch, ia):
    """The AMR fine level (coarse 128^3, ratio 4, 240 patches around 4 shells,
    2^21 points, BSPLINE_3): FP32C and FP64 spread + interp, one launch each
    per level.  Returns the bench record."""
    import numpy as np
    from ibamr_amd import synthetic

    kernel = "BSPLINE_3"
    level, Xn = synthetic.amr_level(coarse=128, ratio=4, npoints=1 << 21, nshell=4, npatches=240)
    n = Xn.shape[0]
    cells = sum(int(np.prod([p.iupper[d] - p.ilower[d] + 1 for d in range(3)])) for p in level.patches)
    X = torch.tensor(Xn, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(99)
    F = torch.randn((n, 3), generator=g, device="cuda", dtype=torch.float64)
    if args.point_order == "regrid":
        # LDataManager's local order after a regrid of this level (patch, cell, Lagrangian index)
        p = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=n)
        Xw = X.clone()
        order, _, _ = p.redistribute(Xw, torch.arange(n, dtype=torch.int32, device="cuda"))
        o = order.long()
        torch.cuda.synchronize()
        p.close()
        X, F = Xw[o].contiguous(), F[o].contiguous()
    print("bench amr: %d patches, %d fine cells, %d points (%s order)" % (len(level.patches), cells, n,
                                                                       args.point_order), file=sys.stderr, flush=True)
    r32 = run_amr_form(torch, ia, ia.FP32C, level, X, F, kernel, args.steps, args.warmup)
    r64 = run_amr_form(torch, ia, ia.FP64, level, X, F, kernel, max(2, args.steps // 2), 1)
    err_f = max(float((a.double() - b).abs().max()) for a, b in zip(r32["f"], r64["f"])) / \
        max(float(b.abs().max()) for b in r64["f"])
    err_u = float((r32["U"].double() - r64["U"]).abs().max() / r64["U"].abs().max())
    # dominant kernel against the HBM roofline: N (8d + 4d) + 4 T bytes per launch,
    # T = distinct grid values in the union of the stencils (fine level index space)
    T = touched_faces(torch, X, 128 * 4, kernel, periodic=False)
    B = n * (8 * 3 + 4 * 3) + 4 * T
    dom = "spread" if r32["kt"]["spread_ms"] >= r32["kt"]["interp_ms"] else "interp"
    t_dom = r32["kt"][dom + "_ms"]
    kname = {"spread": spread_kernel_name(r32["shape"]), "interp": "k_interp3"}[dom]
    achieved = B / (t_dom * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": n / (r32["ms"] * 1e-3), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r32["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "AMR fine level: coarse 128^3 x ratio 4, %d patches (edges 16-64), %d fine cells, "
                               "%d points on 4 ellipsoidal shells, side-centred BSPLINE_3, FP32 compensated"
                               % (len(level.patches), cells, n),
                   "patches": len(level.patches), "fine_cells": cells, "points": n, "kernel": kernel,
                   "point_order": args.point_order, "spread_tiles": r32["info"]["ntiles"],
                   "tile": list(r32["shape"]),
                   "tile_autotune_ms": {"x".join(map(str, h)): t for h, t in r32["tune"].items()},
                   "bins": "%d^3 cells, sorted (CSR: no capacity limit)" % r32["info"]["tile"]},
        "detail": {"spread_kernel_ms": r32["kt"]["spread_ms"], "interp_kernel_ms": r32["kt"]["interp_ms"],
                   "fp64_points_per_s": n / (r64["ms"] * 1e-3), "fp64_spread_kernel_ms": r64["kt"]["spread_ms"],
                   "fp64_interp_kernel_ms": r64["kt"]["interp_ms"], "fp32c_over_fp64": r64["ms"] / r32["ms"],
                   "fp64_tile_autotune_ms": {"x".join(map(str, h)): t for h, t in r64["tune"].items()},
                   "fp32c_rel_linf_err_spread": err_f, "fp32c_rel_linf_err_interp": err_u,
                   "touched_faces": T},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": B},
        "cpu_baseline": None,
    }
    return out


def main_amr(args, torch, ia):
    print(json.dumps(bench_amr(args, torch, ia)), flush=True)


def time_forces(torch, counts, X0, reps=20):
    """ibamr_forces_compute on the shells' points: each shell a closed ring of
    springs (consecutive points) with beams along it, every 4th point a target
    point.  Returns the mean time of one force evaluation (FP32 output)."""
    import numpy as np
    from ibamr_amd import structure

    rng = np.random.default_rng(58)
    m, s, pv = [], [], []
    start = 0
    for c in counts:
        i = np.arange(c)
        m.append(start + i)
        s.append(start + (i + 1) % c)
        pv.append(start + (i - 1) % c)
        start += c
    m, s, pv = map(np.concatenate, (m, s, pv))
    Xh = X0.cpu().numpy()
    tn = np.arange(0, start, 4)
    f = structure.Forces(3, start, springs=(m, s, rng.uniform(1, 2, start), np.full(start, 1e-4)),
                         beams=(pv, m, s, rng.uniform(1e-3, 1e-2, start), np.zeros((start, 3))),
                         targets=(tn, rng.uniform(1, 2, len(tn)), rng.uniform(0, 1, len(tn)), Xh[tn]))
    F = torch.zeros((start, 3), dtype=torch.float32, device="cuda")
    U = torch.zeros((start, 3), dtype=torch.float32, device="cuda")
    f.compute(F, X0, U)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        f.compute(F, X0, U)
    e1.record()
    torch.cuda.synchronize()
    f.close()
    return e0.elapsed_time(e1) / reps


def bench_multi(args, torch, ia):
    """64 ellipsoidal shells on a periodic 384^3 grid, shells 0-31 IB_4 and 32-63
    BSPLINE_3, point counts log-uniform in [2000, 200000]; 100 steps of
    spread -> u = u_bg + alpha f -> ghost fill -> interp -> X += dt U, each step
    one replay of a captured HIP graph.  FP32C and FP64 trajectories compared.
    Returns the bench record."""
    import math

    import numpy as np
    from ibamr_amd import ibstep

    N, nshell, steps = 384, 64, 100
    h = 1.0 / N
    rng = np.random.default_rng(55)
    counts = [int(round(math.exp(v))) for v in rng.uniform(math.log(2000), math.log(200000), nshell)]
    structs = [("IB_4" if i < nshell // 2 else "BSPLINE_3", c) for i, c in enumerate(counts)]
    n = sum(counts)
    level = ia.Level.uniform_periodic(N, 3)
    X0 = ibstep.ellipsoid_shells(nshell, counts, seed=56)
    g = torch.Generator(device="cuda")
    g.manual_seed(57)
    F64 = torch.randn((n, 3), generator=g, device="cuda", dtype=torch.float64)
    X0_nodes = X0  # the shells' node order (the force generator's springs and beams follow it)
    if args.point_order == "regrid":
        # each kernel group's points (one plan each) in LDataManager's post-regrid order
        n4 = sum(c for k, c in structs if k == "IB_4")
        Xa, Fa = regrid_order(torch, ia, X0[:n4], F64[:n4], N, "IB_4")
        Xb, Fb = regrid_order(torch, ia, X0[n4:], F64[n4:], N, "BSPLINE_3")
        X0, F64 = torch.cat([Xa, Xb]).contiguous(), torch.cat([Fa, Fb]).contiguous()
    res = {}
    dt = None
    tunes = {}
    for prec, dt_ in ((ia.FP64, torch.float64), (ia.FP32C, torch.float32)):
        gw = [3] * 3
        S = ibstep.StructureSet(level, structs, prec, alpha=0.05 * h ** 3,
                                u_bg=ibstep.background_velocity(level, gw, dt_))
        X = X0.clone()
        F = F64.to(dt_).contiguous()
        U = torch.zeros((n, 3), dtype=dt_, device="cuda")
        if dt is None:
            # dt from the FP64 velocity of a first eager step; both forms use it
            S.step(X, F, U, 0.0)
            torch.cuda.synchronize()
            dt = 0.25 * h / float(U.abs().max())
            X.copy_(X0)
        S.step(X, F, U, dt)   # warm-up before capture
        torch.cuda.synchronize()
        X.copy_(X0)
        tunes[prec] = S.autotune(X, F)  # spread tile height per kernel group (untimed)
        torch.cuda.synchronize()
        graph = S.capture(X, F, U, dt)
        torch.cuda.synchronize()
        X.copy_(X0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            graph.replay()
        e1.record()
        torch.cuda.synchronize()
        res[prec] = {"ms": e0.elapsed_time(e1) / steps, "X": X.clone(), "dt": dt}
        del S, graph
    drift = float((res[ia.FP32C]["X"] - res[ia.FP64]["X"]).abs().max()) / h
    ms = res[ia.FP32C]["ms"]
    forces_ms = time_forces(torch, counts, X0_nodes)
    # the step against the HBM roofline: per kernel group a spread and an
    # interpolation, each N (8d + 4d) + 4 T bytes (T at the initial positions)
    n4 = sum(c for k, c in structs if k == "IB_4")
    B = 0
    for kern, Xg in (("IB_4", X0[:n4]), ("BSPLINE_3", X0[n4:])):
        B += 2 * (Xg.shape[0] * (8 * 3 + 4 * 3) + 4 * touched_faces(torch, Xg, N, kern))
    achieved = B / (ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": n / (ms * 1e-3), "unit": "points/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "64 ellipsoidal shells (32 IB_4 + 32 BSPLINE_3, %d points) on a periodic 384^3 "
                               "side grid; step = bin + spread + u=u_bg+alpha*f + ghost fill + interp + X+=dtU, "
                               "one HIP-graph replay per step, %d steps, FP32 compensated" % (n, steps),
                   "points": n, "grid": N, "structures": nshell, "point_order": args.point_order,
                   "tile_autotune_ms": {k: {"x".join(map(str, h)): t for h, t in v.items()}
                                        for k, v in tunes[ia.FP32C].items()}},
        "detail": {"fp64_ms_per_step": res[ia.FP64]["ms"], "fp64_points_per_s": n / (res[ia.FP64]["ms"] * 1e-3),
                   "fp64_tile_autotune_ms": {k: {"x".join(map(str, h)): t for h, t in v.items()}
                                             for k, v in tunes[ia.FP64].items()},
                   "fp32c_over_fp64": res[ia.FP64]["ms"] / ms,
                   "dt": res[ia.FP32C]["dt"], "fp32c_vs_fp64_position_drift_cells_after_%d_steps" % steps: drift,
                   "force_generation_ms": forces_ms,
                   "force_generation": "ring springs + beams per shell, a target point every 4th point, "
                                       "timed alone (not in the step above)"},
        "roofline": {"bound": "hbm", "kernel": "step (bin, spread, grid update, ghost fill, interp, Euler update; "
                                              "both kernel groups)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": B},
        "cpu_baseline": None,
    }
    return out


def main_multi(args, torch, ia):
    print(json.dumps(bench_multi(args, torch, ia)), flush=True)


def summary(rec):
    """The figures of a workload's record that the default run reports beside
    its own: throughput, FP32C/FP64, error or drift, roofline fraction."""
    d = rec["detail"]
    out = {"workload": rec["config"]["workload"], "points_per_s": rec["value"], "ms_per_step": rec["ms_per_step"],
           "fp64_points_per_s": d.get("fp64_points_per_s"), "fp32c_over_fp64": d.get("fp32c_over_fp64"),
           "roofline_kernel": rec["roofline"]["kernel"], "roofline_frac": rec["roofline"]["frac"],
           "roofline_achieved_GBs": rec["roofline"]["achieved"]}
    for k, v in d.items():
        if k.startswith("fp32c_rel_linf_err") or k.startswith("fp32c_vs_fp64_position_drift") or k.endswith(
                "kernel_ms"):
            out[k] = v
    if "tile" in rec["config"]:
        out["tile"] = rec["config"]["tile"]
    return out


def main():
    global TILE
    args = parse()
    if args.tile:
        TILE = tuple(int(v) for v in args.tile.split(","))
    if args.gpus != 1:
        print("bench.py measures one GPU only (the library has no multi-GPU path); got --gpus %d" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.copy_probe:
        copy_probe(torch)
        return
    import ibamr_amd as ia

    ia.lib()
    if args.workload == "amr":
        main_amr(args, torch, ia)
        return
    if args.workload == "multi":
        main_multi(args, torch, ia)
        return
    N, n, kernel = args.grid, args.points, args.kernel
    X, F64 = make_points(torch, n, N, seed=2026)
    if args.point_order == "regrid":
        X, F64 = regrid_order(torch, ia, X, F64, N, kernel)
    print("bench: %d^3 %s side, %d points (%s order); FP32C form" % (N, kernel, n, args.point_order),
          file=sys.stderr, flush=True)
    r32 = run_form(torch, ia, ia.FP32C, X, F64, N, kernel, args.steps, args.warmup)
    ms32 = r32["spread_ms"] + r32["interp_ms"]
    T = touched_faces(torch, X, N, kernel)
    d, sx, s32 = 3, 8, 4
    B_spread = n * (sx * d + s32 * d) + s32 * T
    B_interp = n * (sx * d + s32 * d) + s32 * T
    dom = "spread" if r32["spread_kernel_ms"] >= r32["interp_kernel_ms"] else "interp"
    B_dom = B_spread if dom == "spread" else B_interp
    t_dom = r32["spread_kernel_ms"] if dom == "spread" else r32["interp_kernel_ms"]
    achieved = B_dom / (t_dom * 1e-3) / 1e9
    kname = spread_kernel_name(r32["plan"].tile_shape) if dom == "spread" else "k_interp3"
    tag = "%s_%dc_side_%s_%dpts_fp32c" % (kname, N, kernel.lower(), n)
    tr = traffic_from_profile(tag)

    out = {
        "metric": METRIC,
        "value": n / (ms32 * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall_s"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "3-D periodic %d^3 side-centred %s, %d points (half uniform, half in a 2h shell "
                               "r=0.3), FP32 compensated spread+interp" % (N, kernel, n),
                   "grid": N, "points": n, "kernel": kernel, "centering": "side", "precision": "fp32c",
                   "point_order": args.point_order,
                   "tile": list(r32["plan"].tile_shape),
                   "tile_autotune_ms": {"x".join(map(str, h)): t for h, t in r32["tune"].items()},
                   "bins": "%d^3 cells, sorted (CSR: no capacity limit)" % r32["info"]["tile"],
                   "spread_tiles": r32["info"]["ntiles"]},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["bytes_per_dispatch"] if tr else None, "algorithmic_bytes": B_dom,
                     "traffic_source": ("profiles/hbm_traffic_%s.txt" % tag) if tr else None,
                     # what bounds it instead: instruction issue (rocprofv3 counters of this workload)
                     "counters": counters_from_profile(kname, "%dc_side_%s_%dpts_fp32c" % (N, kernel.lower(), n))},
        "detail": {"spread_ms": r32["spread_ms"], "interp_ms": r32["interp_ms"], "bin_ms": r32["bin_ms"],
                   "spread_kernel_ms": r32["spread_kernel_ms"], "interp_kernel_ms": r32["interp_kernel_ms"],
                   "touched_faces": T, "bytes_step": B_spread + B_interp,
                   "step_GBs": (B_spread + B_interp) / (ms32 * 1e-3) / 1e9},
    }
    out["detail"]["redistribute_ms"] = time_redistribute(torch, ia, X, F64, N, kernel)
    out["detail"]["redistribute"] = ("LDataManager redistribution of all points (reversed Lagrangian indices) "
                                     "+ permutation of X and F into the new order, not part of the step")
    if not args.no_fp64:
        print("bench: FP64 form", file=sys.stderr, flush=True)
        r64 = run_form(torch, ia, ia.FP64, X, F64, N, kernel, max(2, args.steps // 2), 1)
        ms64 = r64["spread_ms"] + r64["interp_ms"]
        err_f = max(float((r32["f"][a].double() - r64["f"][a]).abs().max() / r64["f"][a].abs().max()) for a in range(3))
        err_u = float((r32["U"].double() - r64["U"]).abs().max() / r64["U"].abs().max())
        out["detail"].update({"fp64_points_per_s": n / (ms64 * 1e-3), "fp64_spread_ms": r64["spread_ms"],
                              "fp64_interp_ms": r64["interp_ms"], "fp32c_over_fp64": ms64 / ms32,
                              "fp64_tile_autotune_ms": {"x".join(map(str, h)): t for h, t in r64["tune"].items()},
                              "fp32c_rel_linf_err_spread": err_f, "fp32c_rel_linf_err_interp": err_u})
        del r64
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        host = host_oracle()
        Xh, Fh = X.cpu().numpy(), F64.cpu().numpy()
        one = dict(cpu_baseline(Xh, Fh, N, kernel, args.cpu_sample), **host)
        cores = min(args.cpu_cores, os.cpu_count() or 1)
        if cores > 1:
            # the reference's CPU path on the host's cores (one patch per core), the baseline reported;
            # the one-core figure goes in detail
            out["cpu_baseline"] = dict(cpu_baseline_cores(Xh, Fh, N, kernel, args.cpu_sample_cores, cores), **host)
            out["detail"]["cpu_one_core"] = one
        else:
            out["cpu_baseline"] = one
    else:
        out["cpu_baseline"] = None
    if not args.no_others:
        # the AMR-level and multi-structure workloads, summarised beside the headline one
        del r32
        torch.cuda.empty_cache()
        print("bench: AMR level", file=sys.stderr, flush=True)
        out["detail"]["amr_level"] = summary(bench_amr(args, torch, ia))
        torch.cuda.empty_cache()
        print("bench: multi-structure step", file=sys.stderr, flush=True)
        out["detail"]["multi_structure"] = summary(bench_multi(args, torch, ia))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
