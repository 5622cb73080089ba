"""LDataManager::spread with ds_data (LDataManager.cpp:394-464): the points
carry F_ds[k][d] = F[k][d] * ds[k].  The library forms the product while it
puts the points in bin order; the result must be the spread of that product:
FP64 bitwise against the oracle spreading F*ds (numpy forms the same double
products), FP32C bitwise against the library's own spread of F*ds formed on
the device in float (one correctly rounded product either way)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from test_gpu_parity import _arrays, _dev, _oracle_spread  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _level():
    n = 32
    h = 1.0 / n
    boxes = [([4, 4, 4], [11, 19, 11]), ([12, 4, 4], [27, 11, 13]), ([16, 20, 18], [23, 27, 29])]
    patches = [ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]) for lo, hi in boxes]
    level = ia.Level([h] * 3, [0] * 3, [n - 1] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0], patches)
    rng = np.random.default_rng(11)
    X = np.concatenate([np.stack([rng.uniform(lo[d] * h, (hi[d] + 1) * h, 140) for d in range(3)], axis=1)
                        for lo, hi in boxes])
    return level, X, rng


@pytest.mark.parametrize("kernel,centering,depth", [("IB_4", ia.SIDE, 1), ("IB_6", ia.SIDE, 1),
                                                    ("BSPLINE_3", ia.CELL, 2)])
def test_spread_ds_fp64_bitwise_with_oracle(orc, kernel, centering, depth):
    print("case", kernel, centering, depth, flush=True)
    level, X, rng = _level()
    npts = X.shape[0]
    fdepth = 3 if centering == ia.SIDE else depth
    plan = ia.LevelPlan(level, kernel, centering, ia.FP64, depth=depth, max_points=npts)
    F = rng.standard_normal((npts, fdepth))
    ds = rng.uniform(0.5, 2.0, npts)
    Xd, Fd, dsd = _dev(X), _dev(F), _dev(ds)
    plan.bin(Xd)
    init = _arrays(level, plan, rng)
    dv = [_dev(a) for a in init]
    plan.bind(0, dv)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE, ds=dsd)
    ref = _oracle_spread(orc, level, plan, init, X, F * ds[:, None], kernel)
    torch.cuda.synchronize()
    for got, want in zip(dv, ref):
        assert np.array_equal(got.cpu().numpy(), want)
    # the product is not the plain spread of F
    plain = _oracle_spread(orc, level, plan, init, X, F, kernel)
    assert any(not np.array_equal(a.cpu().numpy(), b) for a, b in zip(dv, plain))


def test_spread_ds_fp32c_equals_spread_of_product():
    level, X, rng = _level()
    npts = X.shape[0]
    plan = ia.LevelPlan(level, "IB_6", ia.SIDE, ia.FP32C, max_points=npts)
    Xd = _dev(X)
    F = _dev(rng.standard_normal((npts, 3)).astype(np.float32))
    ds = _dev(rng.uniform(0.5, 2.0, npts).astype(np.float32))
    plan.bin(Xd)
    shapes = [ia.array_shape(level, p, plan.data_gcw, ia.SIDE, axis=a) for p in range(len(level.patches))
              for a in range(3)]
    f1 = [torch.zeros(s, dtype=torch.float32, device=DEV) for s in shapes]
    f2 = [torch.zeros(s, dtype=torch.float32, device=DEV) for s in shapes]
    plan.bind(0, f1)
    plan.bind(1, f2)
    plan.spread(0, F, Xd, mode=ia.ACCUMULATE_INTERIOR, ds=ds)
    plan.spread(1, F * ds[:, None], Xd, mode=ia.ACCUMULATE_INTERIOR)
    torch.cuda.synchronize()
    assert max(float(a.abs().max()) for a in f1) > 0
    for a, b in zip(f1, f2):
        assert torch.equal(a, b)


def test_spread_ds_argument_checks():
    level, X, rng = _level()
    npts = X.shape[0]
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=npts)
    Xd = _dev(X)
    F = _dev(rng.standard_normal((npts, 3)))
    plan.bin(Xd)
    with pytest.raises(ValueError):
        plan.spread(0, F, Xd, ds=torch.ones(npts, dtype=torch.float32, device=DEV))
    with pytest.raises(ValueError):
        plan.spread(0, F, Xd, ds=torch.ones(npts - 1, dtype=torch.float64, device=DEV))
