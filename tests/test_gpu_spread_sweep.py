"""FP32C spread swept along z (tile shapes 8 x 8 x 64/128/256): the wavefront
takes a long z segment through a register window, writing values as the
window passes them.  Against the FP64 spread of the same points and old
values, in every spread mode, within the FP32C tolerance (1e-6 of the largest
value):
  * overwrite: values no stencil reaches are zero (the old values are random,
    so a value the sweep leaves unwritten fails the bar);
  * accumulate (the raw LEInteractor::spread): totals start from the old values;
  * interior (LDataManager::spread): old + sum inside the patch, ghost values
    keep their old bits.
Levels: three patches whose z extents are shorter than a segment (the window
starts and ends inside one segment, and the zero fill covers the rest), a
periodic box (images on both sides), and one tall patch spanning several
segments with points clustered at two heights (windows that jump over empty
layers)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from test_gpu_parity import _arrays, _dev  # noqa: E402

pytestmark = pytest.mark.gpu
SWEEP = [(8, 8, 64), (8, 8, 128), (8, 8, 256)]
MODES = [ia.OVERWRITE, ia.ACCUMULATE, ia.ACCUMULATE_INTERIOR]


def _patches_level():
    n = 32
    h = 1.0 / n
    boxes = [([4, 4, 4], [11, 19, 11]), ([12, 4, 4], [27, 11, 13]), ([16, 20, 18], [23, 27, 29])]
    patches = [ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]) for lo, hi in boxes]
    level = ia.Level([h] * 3, [0] * 3, [n - 1] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0], patches)
    rng = np.random.default_rng(31)
    X = np.concatenate([np.stack([rng.uniform(lo[d] * h, (hi[d] + 1) * h, 150) for d in range(3)], axis=1)
                        for lo, hi in boxes])
    return level, X


def _periodic_level():
    level = ia.Level.uniform_periodic(40, 3)
    rng = np.random.default_rng(32)
    X = rng.uniform(0.0, 1.0, (3000, 3))
    X[:300] = 1.0 - 2e-3 * rng.uniform(0.0, 1.0, (300, 3))   # at the periodic corner
    return level, X


def _tall_level():
    # one 16 x 16 x 300 patch: several segments; points in two clusters of z
    nz = 300
    h = 1.0 / nz
    patch = ia.Patch([0, 0, 0], [15, 15, nz - 1], [0.0] * 3, [16 * h, 16 * h, 1.0])
    level = ia.Level([h] * 3, [0] * 3, [15, 15, nz - 1], [0.0] * 3, [16 * h, 16 * h, 1.0], [0, 0, 0], [patch])
    rng = np.random.default_rng(33)
    X = rng.uniform(0.0, 16 * h, (2000, 3))
    X[:1000, 2] = rng.uniform(20 * h, 40 * h, 1000)
    X[1000:, 2] = rng.uniform(200 * h, 203 * h, 1000)
    return level, X


LEVELS = {"patches": _patches_level, "periodic": _periodic_level, "tall": _tall_level}


def _spread(level, X, F, kernel, centering, precision, old, mode, shape=None):
    plan = ia.LevelPlan(level, kernel, centering, precision, depth=2 if centering == ia.CELL else 1,
                        max_points=X.shape[0])
    if shape is not None:
        plan.tile_shape = shape
    dt = np.float64 if precision == ia.FP64 else np.float32
    Xd = _dev(X)
    plan.bin(Xd)
    dv = [_dev(a.astype(dt)) for a in old]
    plan.bind(0, dv)
    plan.spread(0, _dev(F.astype(dt)), Xd, mode=mode)
    torch.cuda.synchronize()
    return plan, [a.cpu().numpy() for a in dv]


@pytest.mark.parametrize("lvl", sorted(LEVELS))
@pytest.mark.parametrize("kernel,centering", [("IB_6", ia.SIDE), ("IB_4", ia.SIDE), ("BSPLINE_3", ia.CELL)])
def test_sweep_modes_against_fp64(lvl, kernel, centering):
    print("sweep", lvl, kernel, centering, flush=True)
    level, X = LEVELS[lvl]()
    rng = np.random.default_rng(7)
    fdepth = 3 if centering == ia.SIDE else 2
    F = rng.standard_normal((X.shape[0], fdepth)).astype(np.float32).astype(np.float64)
    p0 = ia.LevelPlan(level, kernel, centering, ia.FP64, depth=2 if centering == ia.CELL else 1, max_points=1)
    old = [a.astype(np.float32).astype(np.float64) for a in _arrays(level, p0, rng)]
    for mode in MODES:
        _, want = _spread(level, X, F, kernel, centering, ia.FP64, old, mode)
        scale = max(float(np.abs(a).max()) for a in want)
        for sh in SWEEP:
            print("  mode", mode, "shape", sh, flush=True)
            plan, got = _spread(level, X, F, kernel, centering, ia.FP32C, old, mode, shape=sh)
            assert plan.tile_shape == sh
            err = max(float(np.abs(g.astype(np.float64) - w).max()) for g, w in zip(got, want))
            assert err <= 1e-6 * scale, (mode, sh, err)
            _, again = _spread(level, X, F, kernel, centering, ia.FP32C, old, mode, shape=sh)
            for a, b in zip(got, again):
                assert np.array_equal(a, b), (mode, sh)   # reproducible
            if mode == ia.ACCUMULATE_INTERIOR:
                g = plan.data_gcw
                for a, o in zip(got, old):
                    inner = np.zeros(a.shape, dtype=bool)
                    sl = [slice(g[2 - ax], a.shape[a.ndim - 3 + ax] - g[2 - ax]) for ax in range(3)]
                    inner[(Ellipsis,) + tuple(sl)] = True
                    assert np.array_equal(a[~inner], o.astype(np.float32)[~inner]), sh


def test_sweep_shapes_only_where_admissible():
    level = ia.Level.uniform_periodic(40, 3)
    for kernel in ("IB_4", "IB_6", "BSPLINE_3"):
        p = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP32C, max_points=10)
        for sh in SWEEP:
            p.tile_shape = sh
    # FP64 has no sweep; nor kernels without set-up stencils; nor a reach the window cannot hold
    p64 = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=10)
    pb4 = ia.LevelPlan(level, "BSPLINE_4", ia.SIDE, ia.FP32C, max_points=10)
    pdr = ia.LevelPlan(level, "IB_6", ia.SIDE, ia.FP32C, max_points=10, drift_cells=1)
    for p in (p64, pb4, pdr):
        with pytest.raises(ia.IBAMRError):
            p.tile_shape = (8, 8, 64)
