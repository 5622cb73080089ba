"""Spread tile height and its autotuner (ibamr_plan_set_tile_height /
ibamr_plan_autotune): the height only partitions the work, so FP64 spreads are
bitwise identical at every height (the per-value summation order is fixed by
the bins, and the oracle checks that order elsewhere), FP32C spreads agree to
rounding, and the autotuner keeps the height it measured fastest."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _setup(precision, kernel, centering, n=40, npts=6000, seed=4):
    print("case", precision, kernel, centering, flush=True)
    level = ia.Level.uniform_periodic(n, 3)
    plan = ia.LevelPlan(level, kernel, centering, precision, depth=3 if centering == ia.CELL else 1,
                        max_points=npts)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    X = torch.rand((npts, 3), generator=g, device=DEV, dtype=torch.float64)
    X[: npts // 3] = 0.4 + 0.08 * torch.rand((npts // 3, 3), generator=g, device=DEV, dtype=torch.float64)
    m = min(200, npts // 3)  # near the periodic corner
    X[npts // 3: npts // 3 + m] = 1.0 - 1e-3 * torch.rand((m, 3), generator=g, device=DEV, dtype=torch.float64)
    dt = torch.float64 if precision == ia.FP64 else torch.float32
    F = torch.randn((npts, 3), generator=g, device=DEV, dtype=torch.float64).to(dt)
    narr = 3 if centering == ia.SIDE else 1
    shapes = [ia.array_shape(level, 0, plan.data_gcw, centering, axis=a if centering == ia.SIDE else 0, depth=3)
              for a in range(narr)]
    plan.bin(X)
    return plan, X, F, shapes, dt


def _spread(plan, X, F, shapes, dt):
    f = [torch.zeros(s, dtype=dt, device=DEV) for s in shapes]
    plan.bind(0, f)
    plan.spread(0, F, X, mode=ia.OVERWRITE)
    torch.cuda.synchronize()
    return [a.cpu().numpy() for a in f]


@pytest.mark.parametrize("kernel,centering", [("IB_4", ia.SIDE), ("IB_6", ia.SIDE), ("BSPLINE_3", ia.CELL)])
def test_fp64_bitwise_at_every_height(kernel, centering):
    plan, X, F, shapes, dt = _setup(ia.FP64, kernel, centering)
    assert plan.tile_height == 32
    ref = _spread(plan, X, F, shapes, dt)
    plan.tile_height = 16
    assert plan.tile_height == 16
    got = _spread(plan, X, F, shapes, dt)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kernel", ["IB_6", "BSPLINE_3"])
def test_fp32c_heights_agree_and_autotune(kernel):
    plan, X, F, shapes, dt = _setup(ia.FP32C, kernel, ia.SIDE)
    assert plan.tile_height == 16
    ref = _spread(plan, X, F, shapes, dt)
    scale = max(np.max(np.abs(a)) for a in ref)
    for h in (8, 24, 16):
        plan.tile_height = h
        got = _spread(plan, X, F, shapes, dt)
        err = max(np.max(np.abs(a.astype(np.float64) - b)) for a, b in zip(got, ref)) / scale
        assert err <= 1e-6, (h, err)
    f = [torch.zeros(s, dtype=dt, device=DEV) for s in shapes]
    plan.bind(0, f)
    times = plan.autotune(0, F, X, reps=2)
    print("autotune", times, flush=True)
    assert sorted(times) == [8, 16, 24]
    assert all(t > 0 for t in times.values())
    assert plan.tile_height == min(times, key=times.get)
    # the trials overwrote slot 0 with spreads of F (the last candidate's)
    torch.cuda.synchronize()
    trial = [a.cpu().numpy() for a in f]
    err = max(np.max(np.abs(a.astype(np.float64) - b)) for a, b in zip(trial, ref)) / scale
    assert err <= 1e-6, err
    # at the chosen height the spread is reproducible bit for bit
    once, twice = _spread(plan, X, F, shapes, dt), _spread(plan, X, F, shapes, dt)
    for a, b in zip(once, twice):
        assert np.array_equal(a, b)


def test_height_validation():
    plan, X, F, shapes, dt = _setup(ia.FP64, "IB_4", ia.SIDE, n=16, npts=100)
    for bad in (0, 8, 12, 64):
        with pytest.raises(ia.IBAMRError):
            plan.tile_height = bad
    assert plan.tile_height == 32
    level = ia.Level.uniform_periodic(32, 2)
    p2 = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP32C, max_points=10)
    assert p2.tile_height == 1
    with pytest.raises(ia.IBAMRError):
        p2.tile_height = 16


def test_graph_after_autotune():
    """The autotuner rebuilds the work list (a new device array): a HIP graph
    captured after it replays the spread and interpolation with the chosen tile,
    equal to the eager calls on every replay."""
    plan, X, F, shapes, dt = _setup(ia.FP32C, "IB_4", ia.SIDE)
    f = [torch.zeros(s, dtype=dt, device=DEV) for s in shapes]
    u = [torch.rand(s, dtype=torch.float64, device=DEV).to(dt) for s in shapes]
    plan.bind(0, f)
    plan.bind(1, u)
    plan.autotune(0, F, X, reps=1)
    U = torch.zeros((X.shape[0], 3), dtype=dt, device=DEV)
    plan.spread(0, F, X, mode=ia.OVERWRITE)
    plan.interp(1, U, None)
    torch.cuda.synchronize()
    f_ref, U_ref = [a.clone() for a in f], U.clone()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            plan.spread(0, F, X, mode=ia.OVERWRITE, stream=side)
            plan.interp(1, U, None, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(3):
        for a in f:
            a.fill_(7.0)
        U.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(f, f_ref):
            assert torch.equal(a, b), rep
        assert torch.equal(U, U_ref), rep
