"""GPU parity on a synthetic AMR fine level (many patches of different sizes,
not periodic): FP64 spread and interp bitwise against the oracle run patch by
patch; FP32C within 1e-6 of FP64.  Coarse 16^3 refined by 4 (fine 64^3)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from ibamr_amd import synthetic  # noqa: E402
from test_gpu_parity import _oracle_interp, _oracle_spread  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.mark.parametrize("kernel", ["BSPLINE_3", "IB_4", "BSPLINE_5", "BSPLINE_6"])
def test_amr_level_fp64_bitwise(orc, kernel):
    level, X = synthetic.amr_level(coarse=16, ratio=4, npoints=4000, nshell=4, seed=5, block=16)
    assert len(level.patches) > 8
    rng = np.random.default_rng(17)
    F = rng.standard_normal(X.shape)
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=X.shape[0])
    arrays = []
    for p in range(len(level.patches)):
        for a in range(3):
            arrays.append(rng.uniform(-1, 1, ia.array_shape(level, p, plan.data_gcw, ia.SIDE, axis=a)))
    print("amr level", kernel, len(level.patches), "patches", flush=True)
    Xd, Fd = _dev(X), _dev(F)
    plan.bin(Xd)
    dev = [_dev(a) for a in arrays]
    plan.bind(0, dev)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE)
    torch.cuda.synchronize()
    ref = _oracle_spread(orc, level, plan, arrays, X, F, kernel)
    for i in range(len(arrays)):
        assert np.array_equal(dev[i].cpu().numpy(), ref[i]), i
    plan.bind(1, [_dev(a) for a in arrays])
    U = torch.zeros((X.shape[0], 3), dtype=torch.float64, device="cuda")
    plan.interp(1, U, Xd)
    torch.cuda.synchronize()
    Uref = _oracle_interp(orc, level, plan, arrays, X, kernel)
    assert not np.isnan(Uref).any()  # every point lies in a patch interior
    assert np.array_equal(U.cpu().numpy(), Uref)


def test_amr_level_fp32c_error(orc):
    kernel = "BSPLINE_3"
    level, X = synthetic.amr_level(coarse=16, ratio=4, npoints=4000, nshell=4, seed=6, block=16)
    rng = np.random.default_rng(18)
    F = rng.standard_normal(X.shape)
    arrays = []
    for p in range(len(level.patches)):
        for a in range(3):
            arrays.append(rng.uniform(-1, 1, ia.array_shape(level, p, [3] * 3, ia.SIDE, axis=a)))
    out = {}
    for prec, dt in ((ia.FP64, torch.float64), (ia.FP32C, torch.float32)):
        plan = ia.LevelPlan(level, kernel, ia.SIDE, prec, max_points=X.shape[0])
        Xd = _dev(X)
        plan.bin(Xd)
        f = [torch.zeros(a.shape, dtype=dt, device="cuda") for a in arrays]
        plan.bind(0, f)
        plan.spread(0, _dev(F).to(dt), Xd, mode=ia.OVERWRITE)
        plan.bind(1, [_dev(a).to(dt) for a in arrays])
        U = torch.zeros((X.shape[0], 3), dtype=dt, device="cuda")
        plan.interp(1, U, Xd)
        torch.cuda.synchronize()
        out[prec] = ([t.double().cpu().numpy() for t in f], U.double().cpu().numpy())
    f64, U64 = out[ia.FP64]
    f32, U32 = out[ia.FP32C]
    scale = max(np.abs(a).max() for a in f64)
    assert max(np.abs(a - b).max() for a, b in zip(f32, f64)) <= 1e-6 * scale
    assert np.abs(U32 - U64).max() <= 1e-6 * np.abs(U64).max()


def test_amr_level_full_size(orc):
    """The full AMR workload: coarse 128^3 refined by 4, 240 patches with edges
    drawn from {16, 24, 32, 48, 64} around 4 ellipsoidal shells, 2^21 points,
    side-centred BSPLINE_3.  FP64 spread (accumulate onto a random field) and
    interp bitwise against the oracle on every patch; FP32C spread and interp
    within 1e-6 (relative L-infinity) of FP64."""
    kernel = "BSPLINE_3"
    level, X = synthetic.amr_level(coarse=128, ratio=4, npoints=1 << 21, nshell=4, npatches=240)
    assert len(level.patches) == 240
    print("amr full size:", len(level.patches), "patches", X.shape[0], "points", flush=True)
    rng = np.random.default_rng(19)
    F = rng.standard_normal(X.shape)
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=X.shape[0])
    arrays = [rng.uniform(-1, 1, ia.array_shape(level, p, plan.data_gcw, ia.SIDE, axis=a))
              for p in range(len(level.patches)) for a in range(3)]
    Xd, Fd = _dev(X), _dev(F)
    plan.bin(Xd)
    dev = [_dev(a) for a in arrays]
    plan.bind(0, dev)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE)
    plan.bind(1, [_dev(a) for a in arrays])
    U = torch.zeros((X.shape[0], 3), dtype=torch.float64, device="cuda")
    plan.interp(1, U, Xd)
    torch.cuda.synchronize()
    ref = _oracle_spread(orc, level, plan, arrays, X, F, kernel)
    for i in range(len(arrays)):
        assert np.array_equal(dev[i].cpu().numpy(), ref[i]), i
    Uref = _oracle_interp(orc, level, plan, arrays, X, kernel)
    assert not np.isnan(Uref).any()
    assert np.array_equal(U.cpu().numpy(), Uref)
    # FP32C against the FP64 results just checked: spread from zero, interp of the random field
    plan32 = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP32C, max_points=X.shape[0])
    plan32.bin(Xd)
    f32 = [torch.zeros(a.shape, dtype=torch.float32, device="cuda") for a in arrays]
    plan32.bind(0, f32)
    plan32.spread(0, Fd.float(), Xd, mode=ia.OVERWRITE)
    plan32.bind(1, [_dev(a).float() for a in arrays])
    U32 = torch.zeros((X.shape[0], 3), dtype=torch.float32, device="cuda")
    plan32.interp(1, U32, Xd)
    f64 = [torch.zeros(a.shape, dtype=torch.float64, device="cuda") for a in arrays]
    plan.bind(2, f64)
    plan.spread(2, Fd, Xd, mode=ia.OVERWRITE)
    torch.cuda.synchronize()
    scale = max(float(a.abs().max()) for a in f64)
    err = max(float((a.double() - b).abs().max()) for a, b in zip(f32, f64))
    print("amr full size FP32C errors: spread %.3g interp %.3g" % (
        err / scale, float((U32.double() - U).abs().max() / U.abs().max())), flush=True)
    assert err <= 1e-6 * scale
    assert float((U32.double() - U).abs().max()) <= 1e-6 * float(U.abs().max())
