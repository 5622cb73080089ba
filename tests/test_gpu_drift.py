"""Points that moved after their binning (drift_cells).  LDataManager keeps each
point in the index sets of the last regrid and spreads / interpolates it at its
current position (LDataManager.cpp:552-667, 699-813): between regrids a point's
stencil follows its new position while its patch and periodic images follow
its old cell.  A plan created with drift_cells = d takes stencils up to d cells
away from the binned cell.  Against the oracle given the old cells and the new
positions: FP64 bitwise (spread in accumulate mode, interpolation); FP32C with
every built tile shape (the z-sweep segments included where the window holds
the wider reach) within the FP32C tolerance of the FP64 result."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from test_gpu_parity import _arrays, _dev, _periodic_fill_np  # noqa: E402

pytestmark = pytest.mark.gpu


def _periodic():
    n = 32
    level = ia.Level.uniform_periodic(n, 3)
    rng = np.random.default_rng(61)
    X0 = rng.uniform(0.0, 1.0, (3000, 3))
    X0[:300] = 1.0 - 2e-3 * rng.uniform(0.0, 1.0, (300, 3))   # at the periodic corner
    return level, X0, rng, True


def _patches():
    n = 32
    h = 1.0 / n
    boxes = [([4, 4, 4], [11, 19, 11]), ([12, 4, 4], [27, 11, 13]), ([16, 20, 18], [23, 27, 29])]
    patches = [ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]) for lo, hi in boxes]
    level = ia.Level([h] * 3, [0] * 3, [n - 1] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0], patches)
    rng = np.random.default_rng(62)
    X0 = np.concatenate([np.stack([rng.uniform(lo[d] * h, (hi[d] + 1) * h, 200) for d in range(3)], axis=1)
                         for lo, hi in boxes])
    return level, X0, rng, False


def _moved(level, X0, rng, drift, periodic):
    h = level.dx[0]
    X1 = X0 + rng.uniform(-0.99, 0.99, X0.shape) * drift * h
    if periodic:
        return X1   # images follow the old cell; positions may leave the box by less than a cell
    return np.clip(X1, 0.0, np.nextafter(1.0, 0.0))


def _oracle(orc, level, plan, kernel, init, u, X0, X1, F):
    """Spread (accumulate into init) and interp (from u) with the old cells and the new positions."""
    cells0 = orc.get_cell_index(X0, level.domain_x_lower, level.domain_x_upper, level.dx, level.domain_lower,
                                level.domain_upper)
    ranks = orc.bin_ranks(cells0, level.domain_lower, [8] * 3)
    igcw = ia.get_minimum_ghost_width(kernel)
    f = [a.copy() for a in init]
    U = np.full((X1.shape[0], 3), np.nan)
    for pi, P in enumerate(level.patches):
        orc.spread_patch(kernel, True, level.dx, P.x_lower, P.ilower, P.iupper, plan.data_gcw, f[3 * pi:3 * pi + 3],
                         X1, F, cells0, igcw, level.periodic_shift, level.domain_lower, ranks=ranks)
        idx, sh = orc.build_patch_entries(cells0, P.ilower, P.iupper, 0, level.periodic_shift, level.domain_lower,
                                          level.dx, [8] * 3, interior_only=True)
        if len(idx):
            V = orc.interp_side(kernel, level.dx, P.x_lower, P.ilower, P.iupper, plan.data_gcw, u[3 * pi:3 * pi + 3],
                                X1, indices=idx, Xshift=sh)
            U[idx] = V[idx]
    return f, U


def _run(level, kernel, precision, drift, X0, X1, F, init, u, shape=None):
    plan = ia.LevelPlan(level, kernel, ia.SIDE, precision, max_points=X0.shape[0], drift_cells=drift)
    if shape is not None:
        plan.tile_shape = shape
    dt = np.float64 if precision == ia.FP64 else np.float32
    plan.bin(_dev(X0))
    f = [_dev(a.astype(dt)) for a in init]
    plan.bind(0, f)
    plan.bind(1, [_dev(a.astype(dt)) for a in u])
    X1d = _dev(X1)
    plan.spread(0, _dev(F.astype(dt)), X1d, mode=ia.ACCUMULATE)
    U = torch.full((X1.shape[0], 3), float("nan"), dtype=torch.float64 if dt == np.float64 else torch.float32,
                   device="cuda")
    plan.interp(1, U, X1d)
    torch.cuda.synchronize()
    return plan, [a.cpu().numpy() for a in f], U.cpu().numpy()


CASES = [("IB_4", 1), ("IB_4", 2), ("IB_6", 1), ("BSPLINE_3", 1)]


@pytest.mark.parametrize("lvl", ["periodic", "patches"])
@pytest.mark.parametrize("kernel,drift", CASES)
def test_drift_fp64_bitwise_and_fp32c(orc, lvl, kernel, drift):
    print("drift", lvl, kernel, drift, flush=True)
    level, X0, rng, periodic = (_periodic if lvl == "periodic" else _patches)()
    X1 = _moved(level, X0, rng, drift, periodic)
    F = rng.standard_normal(X0.shape).astype(np.float32).astype(np.float64)
    p0 = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=1)
    init = [a.astype(np.float32).astype(np.float64) for a in _arrays(level, p0, rng)]
    u = [a.astype(np.float32).astype(np.float64) for a in _arrays(level, p0, rng)]
    if periodic:
        u = [_periodic_fill_np(a, p0.data_gcw[0], 32) for a in u]
    plan, f64, U64 = _run(level, kernel, ia.FP64, drift, X0, X1, F, init, u)
    fo, Uo = _oracle(orc, level, plan, kernel, init, u, X0, X1, F)
    for a, b in zip(f64, fo):
        assert np.array_equal(a, b), np.abs(a - b).max()
    assert np.array_equal(U64, Uo), np.nanmax(np.abs(U64 - Uo))
    # FP32C: every shape this plan builds
    p32 = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP32C, max_points=X0.shape[0], drift_cells=drift)
    cand = [(8, 8, 16), (8, 8, 8), (8, 8, 24), (16, 4, 16), (8, 8, 64), (8, 8, 128), (8, 8, 256)]
    built = []
    for sh in cand:
        try:
            p32.tile_shape = sh
            built.append(sh)
        except ia.IBAMRError:
            pass
    assert (8, 8, 16) in built
    sweep = bool([s for s in built if s[2] >= 48])
    assert sweep == (8 + 2 * (ia.get_minimum_ghost_width(kernel) + drift) <= 16)
    scale = max(np.abs(a).max() for a in f64)
    uscale = np.nanmax(np.abs(U64))
    for sh in built:
        _, f32, U32 = _run(level, kernel, ia.FP32C, drift, X0, X1, F, init, u, shape=sh)
        err = max(np.abs(a.astype(np.float64) - b).max() for a, b in zip(f32, f64))
        assert err <= 1e-6 * scale, (sh, err)
        ok = ~np.isnan(U64)
        assert np.array_equal(np.isnan(U32), ~ok)
        assert np.abs(U32[ok].astype(np.float64) - U64[ok]).max() <= 1e-6 * uscale, sh
