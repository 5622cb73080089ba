"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

FP64 plans must match the oracle bitwise:
  * interp: every point, every axis/component (same per-point arithmetic);
  * spread: every grid value of every ghosted array, starting from the same
    non-zero field, against the oracle run over the patch's entry list
    (points + periodic images, LIndexSetData.cpp:55-150) in the order the
    library sums contributions (oracle.spread_patch: bin, image, 64-point
    chunk, stencil z-start, point).
FP32C plans: relative L-infinity error against the FP64 oracle <= 1e-6 (the
bar of the precision form).
"""
import os
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL32 = 1e-6


def _points(ndim, n, npts, seed, crowd=True):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.0, 1.0, (npts, ndim))
    if crowd:
        h = 1.0 / n
        q = npts // 4
        X[:q] = rng.uniform(0.0, 2.5 * h, (q, ndim))              # near the periodic corner
        X[q:2 * q] = 1.0 - rng.uniform(1e-12, 2.5 * h, (q, ndim))
        X[2 * q:2 * q + 8] = np.array([[0.5] * ndim]) + (np.arange(8)[:, None] % 2) * h * 0.5  # cell faces/centres
    return X


def _periodic_fill_np(u, gw, n):
    out = u.copy()
    for ax in range(out.ndim - 1, -1, -1):
        if out.shape[ax] < n:  # depth axis
            continue
        idx = (np.arange(out.shape[ax]) - gw) % n + gw
        out = np.take(out, idx, axis=ax)
    return out


def _arrays(level, plan, rng, dtype=np.float64, fill="random"):
    out = []
    for p in range(len(level.patches)):
        for a in range(plan.narr):
            shp = ia.array_shape(level, p, plan.data_gcw, plan.centering, axis=a, depth=plan.depth)
            v = rng.uniform(-1.0, 1.0, shp) if fill == "random" else np.zeros(shp)
            out.append(np.ascontiguousarray(v.astype(dtype)))
    return out


def _tile(ndim):
    return 8 if ndim == 3 else 16


def _oracle_spread(orc, level, plan, arrays_np, X, F, kernel):
    """LDataManager::spread over all patches of the level with the oracle, in the
    library's summation order (the FP64 comparison is bitwise)."""
    ndim = level.ndim
    igcw = ia.get_minimum_ghost_width(kernel)
    cells = orc.get_cell_index(X, level.domain_x_lower, level.domain_x_upper, level.dx, level.domain_lower,
                               level.domain_upper)
    out = [a.copy() for a in arrays_np]
    # each point's rank in its bin: one computation for all the level's patches
    ranks = orc.bin_ranks(cells, level.domain_lower, [8] * 3) if ndim == 3 else None

    def one(pi):
        P = level.patches[pi]
        q = out[pi * ndim:(pi + 1) * ndim] if plan.centering == ia.SIDE else out[pi]
        orc.spread_patch(kernel, plan.centering == ia.SIDE, level.dx, P.x_lower, P.ilower, P.iupper, plan.data_gcw,
                         q, X, F, cells, igcw, level.periodic_shift, level.domain_lower, depth=plan.depth,
                         ranks=ranks)

    _over_patches(one, len(level.patches))  # every patch writes only its own arrays
    return out


def _over_patches(fn, npatches):
    """fn(p) for every patch on host threads (the oracle's C calls release the GIL)."""
    if npatches < 4:
        for p in range(npatches):
            fn(p)
        return
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(fn, range(npatches)))


def _oracle_interp(orc, level, plan, arrays_np, X, kernel):
    ndim = level.ndim
    cells = orc.get_cell_index(X, level.domain_x_lower, level.domain_x_upper, level.dx, level.domain_lower,
                               level.domain_upper)
    fdepth = ndim if plan.centering == ia.SIDE else plan.depth
    U = np.full((X.shape[0], fdepth), np.nan)

    def one(pi):
        P = level.patches[pi]
        idx, sh = orc.build_patch_entries(cells, P.ilower, P.iupper, 0, level.periodic_shift, level.domain_lower,
                                          level.dx, [_tile(ndim)] * ndim, interior_only=True)
        if len(idx) == 0:
            return
        if plan.centering == ia.SIDE:
            q = arrays_np[pi * ndim:(pi + 1) * ndim]
            V = orc.interp_side(kernel, level.dx, P.x_lower, P.ilower, P.iupper, plan.data_gcw, q, X, indices=idx,
                                Xshift=sh)
        else:
            V = orc.interp(kernel, level.dx, P.x_lower, P.ilower, P.iupper, plan.data_gcw, arrays_np[pi], X,
                           indices=idx, Xshift=sh, depth=plan.depth)
        U[idx] = V[idx]  # the patches' interior points are disjoint

    _over_patches(one, len(level.patches))
    return U


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


CASES = [
    # ndim, n, kernel, centering, depth, npts
    (3, 16, "IB_4", ia.SIDE, 1, 600),
    (3, 16, "IB_6", ia.SIDE, 1, 600),
    (3, 16, "BSPLINE_3", ia.SIDE, 1, 600),
    (3, 12, "IB_6", ia.CELL, 2, 400),
    (3, 8, "IB_4", ia.CELL, 1, 300),
    (2, 32, "IB_4", ia.SIDE, 1, 500),
    (2, 32, "IB_6", ia.SIDE, 1, 500),
    (2, 24, "BSPLINE_3", ia.CELL, 3, 400),
    (3, 16, "BSPLINE_4", ia.SIDE, 1, 600),
    (3, 16, "BSPLINE_5", ia.SIDE, 1, 600),
    (3, 16, "BSPLINE_6", ia.SIDE, 1, 600),
    (3, 12, "BSPLINE_5", ia.CELL, 1, 300),
    (2, 32, "BSPLINE_4", ia.SIDE, 1, 500),
    (2, 24, "BSPLINE_6", ia.CELL, 2, 400),
]


def _ids(c):
    return "%dd-n%d-%s-%s-d%d" % (c[0], c[1], c[2], "side" if c[3] == ia.SIDE else "cell", c[4])


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_fp64_periodic_bitwise(orc, case):
    ndim, n, kernel, centering, depth, npts = case
    print("case", _ids(case), flush=True)
    rng = np.random.default_rng(zlib.crc32(_ids(case).encode()))
    level = ia.Level.uniform_periodic(n, ndim)
    plan = ia.LevelPlan(level, kernel, centering, ia.FP64, depth=depth, max_points=npts)
    X = _points(ndim, n, npts, 7)
    fdepth = ndim if centering == ia.SIDE else depth
    F = rng.standard_normal((npts, fdepth))
    Xd, Fd = _dev(X), _dev(F)
    plan.bin(Xd)
    torch.cuda.synchronize()

    # --- spread (accumulate into a non-zero field)
    init = _arrays(level, plan, rng)
    dev_arrays = [_dev(a) for a in init]
    plan.bind(0, dev_arrays)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE)
    torch.cuda.synchronize()
    ref = _oracle_spread(orc, level, plan, init, X, F, kernel)
    for got, want in zip(dev_arrays, ref):
        g = got.cpu().numpy()
        assert np.array_equal(g, want), np.max(np.abs(g - want))

    # --- spread twice from the same field: bitwise reproducible
    again = [_dev(a) for a in init]
    plan.bind(1, again)
    plan.spread(1, Fd, Xd, mode=ia.ACCUMULATE)
    torch.cuda.synchronize()
    for a, b in zip(dev_arrays, again):
        assert torch.equal(a, b)

    # --- overwrite mode = zero field + spread
    ow = [_dev(a) for a in init]
    plan.bind(2, ow)
    plan.spread(2, Fd, Xd, mode=ia.OVERWRITE)
    zero = [np.zeros_like(a) for a in init]
    ref0 = _oracle_spread(orc, level, plan, zero, X, F, kernel)
    torch.cuda.synchronize()
    for got, want in zip(ow, ref0):
        assert np.array_equal(got.cpu().numpy(), want)

    # --- periodic ghost fill on the GPU, then interp
    u_np = []
    for a in init:
        gw = plan.data_gcw[0]
        u_np.append(_periodic_fill_np(a, gw, n))
    u_dev = [_dev(a) for a in init]
    plan.bind(3, u_dev)
    plan.fill_periodic_ghosts(3)
    torch.cuda.synchronize()
    for got, want in zip(u_dev, u_np):
        assert np.array_equal(got.cpu().numpy(), want)
    U = torch.full((npts, fdepth), float("nan"), dtype=torch.float64, device=DEV)
    plan.interp(3, U, Xd)
    torch.cuda.synchronize()
    Uref = _oracle_interp(orc, level, plan, u_np, X, kernel)
    assert np.array_equal(U.cpu().numpy(), Uref), np.nanmax(np.abs(U.cpu().numpy() - Uref))


FP32_CASES = CASES[:3] + CASES[5:7] + CASES[8:11] + CASES[12:13]


@pytest.mark.parametrize("case", FP32_CASES, ids=[_ids(c) for c in FP32_CASES])
def test_fp32c_error(orc, case):
    ndim, n, kernel, centering, depth, npts = case
    print("case", _ids(case), flush=True)
    rng = np.random.default_rng(11)
    level = ia.Level.uniform_periodic(n, ndim)
    plan = ia.LevelPlan(level, kernel, centering, ia.FP32C, depth=depth, max_points=npts)
    X = _points(ndim, n, npts, 5)
    fdepth = ndim
    F = rng.standard_normal((npts, fdepth))
    Xd, Fd = _dev(X), _dev(F.astype(np.float32))
    plan.bin(Xd)
    zero = _arrays(level, plan, rng, fill="zero")
    f_dev = [_dev(a.astype(np.float32)) for a in zero]
    plan.bind(0, f_dev)
    plan.spread(0, Fd, Xd, mode=ia.OVERWRITE)
    ref = _oracle_spread(orc, level, plan, zero, X, F.astype(np.float32).astype(np.float64), kernel)
    torch.cuda.synchronize()
    for got, want in zip(f_dev, ref):
        err = np.max(np.abs(got.cpu().numpy().astype(np.float64) - want)) / np.max(np.abs(want))
        assert err <= TOL32, err
    u = [_periodic_fill_np(a, plan.data_gcw[0], n).astype(np.float32) for a in _arrays(level, plan, rng)]
    u_dev = [_dev(a) for a in u]
    plan.bind(1, u_dev)
    U = torch.zeros((npts, fdepth), dtype=torch.float32, device=DEV)
    plan.interp(1, U, Xd)
    Uref = _oracle_interp(orc, level, plan, [a.astype(np.float64) for a in u], X, kernel)
    torch.cuda.synchronize()
    err = np.max(np.abs(U.cpu().numpy().astype(np.float64) - Uref)) / np.max(np.abs(Uref))
    assert err <= TOL32, err


def test_multi_patch_level_fp64(orc):
    """A non-periodic level of three patches (one of them touching a neighbour):
    spread over each patch's index ghost box (points of neighbouring patches
    included, LDataManager.cpp:623-658), interp from the owning patch."""
    ndim, n, kernel = 3, 32, "IB_4"
    h = 1.0 / n
    boxes = [([4, 4, 4], [11, 19, 11]), ([12, 4, 4], [27, 11, 13]), ([16, 20, 18], [23, 27, 29])]
    patches = [ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]) for lo, hi in boxes]
    level = ia.Level([h] * 3, [0] * 3, [n - 1] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0], patches)
    rng = np.random.default_rng(3)
    pts = []
    for lo, hi in boxes:
        m = 150
        pts.append(np.stack([rng.uniform(lo[d] * h, (hi[d] + 1) * h, m) for d in range(3)], axis=1))
    X = np.concatenate(pts)
    npts = X.shape[0]
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=npts)
    F = rng.standard_normal((npts, 3))
    Xd, Fd = _dev(X), _dev(F)
    plan.bin(Xd)
    init = _arrays(level, plan, rng)
    dv = [_dev(a) for a in init]
    plan.bind(0, dv)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE)
    ref = _oracle_spread(orc, level, plan, init, X, F, kernel)
    torch.cuda.synchronize()
    for got, want in zip(dv, ref):
        assert np.array_equal(got.cpu().numpy(), want)
    uv = [_dev(a) for a in init]
    plan.bind(1, uv)
    U = torch.full((npts, 3), float("nan"), dtype=torch.float64, device=DEV)
    plan.interp(1, U, Xd)
    Uref = _oracle_interp(orc, level, plan, init, X, kernel)
    torch.cuda.synchronize()
    assert np.array_equal(U.cpu().numpy(), Uref)
    # X = NULL: the positions the last call put in bin order are reused
    U2 = torch.full((npts, 3), float("nan"), dtype=torch.float64, device=DEV)
    plan.interp(1, U2, None)
    torch.cuda.synchronize()
    assert torch.equal(U2, U)
    plan.bin(Xd)  # a new binning invalidates them
    with pytest.raises(ia.IBAMRError):
        plan.interp(1, U2, None)


def test_out_of_domain_reported(orc):
    level = ia.Level.uniform_periodic(8, 3)
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=4)
    X = _dev(np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.1, np.nan, 0.2], [0.2, 0.2, 0.2]]))
    with pytest.raises(ia.IBAMRError) as e:
        plan.bin(X)
    assert e.value.code == ia.IBAMR_ERR_OUT_OF_DOMAIN


@pytest.mark.parametrize("precision", [ia.FP64, ia.FP32C])
def test_spread_interp_bitwise_reproducible(precision):
    """No floating-point atomics: two runs of the same spread and interp give
    identical bits, also after re-binning (a fresh binning gives the same order).
    The interpolation gives the same bits whether it evaluates the stencils
    itself (X passed) or reads those the spread set-up computed (X = None)."""
    n, npts, kernel = 32, 5000, "IB_6"
    dt = torch.float64 if precision == ia.FP64 else torch.float32
    level = ia.Level.uniform_periodic(n, 3)
    plan = ia.LevelPlan(level, kernel, ia.SIDE, precision, max_points=npts)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    X = torch.rand((npts, 3), generator=g, device="cuda", dtype=torch.float64)
    X[: npts // 2] = 0.5 + 0.05 * torch.rand((npts // 2, 3), generator=g, device="cuda", dtype=torch.float64)
    F = torch.randn((npts, 3), generator=g, device="cuda", dtype=torch.float64).to(dt)
    shapes = [ia.array_shape(level, 0, plan.data_gcw, ia.SIDE, axis=a) for a in range(3)]
    u = [torch.rand(s, generator=g, device="cuda", dtype=torch.float64).to(dt) for s in shapes]
    plan.bind(1, u)
    outs = []
    for rep in range(2):
        plan.bin(X)
        f = [torch.zeros(s, dtype=dt, device="cuda") for s in shapes]
        plan.bind(0, f)
        plan.spread(0, F, X, mode=ia.OVERWRITE)
        U = torch.zeros((npts, 3), dtype=dt, device="cuda")
        plan.interp(1, U, X)
        plan.spread(0, F, X, mode=ia.OVERWRITE)
        U2 = torch.zeros((npts, 3), dtype=dt, device="cuda")
        plan.interp(1, U2, None)
        torch.cuda.synchronize()
        assert torch.equal(U2, U)
        outs.append((f, U))
    for a in range(3):
        assert torch.equal(outs[0][0][a], outs[1][0][a]), a
    assert torch.equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("precision", [ia.FP64, ia.FP32C])
def test_shifted_index_space(precision):
    """The same periodic level with its index space starting at (5, -3, 7)
    instead of 0 (SAMRAI boxes may start anywhere): spread and interp give the
    same bits, in both forms (FP32C places stencils in the level's frame, lower
    index included)."""
    n, npts, kernel = 24, 4000, "IB_6"
    lo = [5, -3, 7]
    level0 = ia.Level.uniform_periodic(n, 3)
    p = ia.Patch(lo, [l + n - 1 for l in lo], [0.0] * 3, [1.0] * 3)
    level1 = ia.Level([1.0 / n] * 3, lo, [l + n - 1 for l in lo], [0.0] * 3, [1.0] * 3, [n] * 3, [p])
    dt = torch.float64 if precision == ia.FP64 else torch.float32
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    X = torch.rand((npts, 3), generator=g, device=DEV, dtype=torch.float64)
    X[:500] = 1.0 - 0.05 * torch.rand((500, 3), generator=g, device=DEV, dtype=torch.float64)  # periodic images
    F = torch.randn((npts, 3), generator=g, device=DEV, dtype=torch.float64)
    out = []
    for level in (level0, level1):
        plan = ia.LevelPlan(level, kernel, ia.SIDE, precision, max_points=npts)
        shapes = [ia.array_shape(level, 0, plan.data_gcw, ia.SIDE, axis=a) for a in range(3)]
        f = [torch.zeros(s, dtype=dt, device=DEV) for s in shapes]
        u = [torch.sin(torch.arange(int(np.prod(s)), device=DEV, dtype=torch.float64) * 0.01).reshape(s).to(dt)
             for s in shapes]
        plan.bin(X)
        plan.bind(0, f)
        plan.bind(1, u)
        plan.spread(0, F.to(dt), X, mode=ia.OVERWRITE)
        U = torch.zeros((npts, 3), dtype=dt, device=DEV)
        plan.interp(1, U, None)
        torch.cuda.synchronize()
        out.append(([a.cpu().numpy() for a in f], U.cpu().numpy()))
        plan.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("centering", [ia.SIDE, ia.CELL])
def test_interior_mode_adds_old_value_once(orc, centering):
    """IBAMR_SPREAD_ACCUMULATE_INTERIOR is LDataManager::spread's zero, spread,
    add (LDataManager.cpp:594, 661-663): every interior value becomes
    old + (sum of its contributions from zero), bit for bit, and ghost values
    keep their old values."""
    ndim, n, kernel = 3, 32, "IB_4"
    h = 1.0 / n
    boxes = [([4, 4, 4], [11, 19, 11]), ([12, 4, 4], [27, 11, 13]), ([16, 20, 18], [23, 27, 29])]
    patches = [ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]) for lo, hi in boxes]
    level = ia.Level([h] * 3, [0] * 3, [n - 1] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0], patches)
    rng = np.random.default_rng(23)
    X = np.concatenate([np.stack([rng.uniform(lo[d] * h, (hi[d] + 1) * h, 120) for d in range(3)], axis=1)
                        for lo, hi in boxes])
    depth = 1 if centering == ia.SIDE else 2
    plan = ia.LevelPlan(level, kernel, centering, ia.FP64, depth=depth, max_points=X.shape[0])
    F = rng.standard_normal((X.shape[0], 3 if centering == ia.SIDE else depth))
    Xd, Fd = _dev(X), _dev(F)
    plan.bin(Xd)
    base = _arrays(level, plan, rng)
    dv = [_dev(a) for a in base]
    plan.bind(0, dv)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE_INTERIOR)
    zero = _oracle_spread(orc, level, plan, [np.zeros_like(a) for a in base], X, F, kernel)
    torch.cuda.synchronize()
    g = plan.data_gcw
    for i, (got, old, s) in enumerate(zip(dv, base, zero)):
        got = got.cpu().numpy()
        inner = np.zeros(got.shape, dtype=bool)
        # array index order (depth,) z, y, x; the interior (side box along the array's axis)
        sl = [slice(g[2 - ax], got.shape[got.ndim - 3 + ax] - g[2 - ax]) for ax in range(3)]
        inner[(Ellipsis,) + tuple(sl)] = True
        assert np.array_equal(got[inner], (old + s)[inner]), i
        assert np.array_equal(got[~inner], old[~inner]), i
