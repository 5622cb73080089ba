"""Spread tile shapes and their autotuner (ibamr_plan_set_tile_shape /
ibamr_plan_autotune).  A shape only partitions the work: FP64 spreads are
bitwise identical with every shape (the per-value summation order is fixed by
the bins, and the oracle checks that order elsewhere).  FP32C spreads are
bitwise reproducible for a given shape; across shapes (which group the
points into different 64-point chunks, each summed in FP32 and folded into
FP64 totals) and across point orders they agree within the FP32C tolerance.
Work lists of all shapes are built with the plan, so a graph captured with
one shape stays valid after another is chosen."""
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


FP64_SHAPES = [(8, 8, 32), (8, 8, 16), (16, 4, 32)]


@pytest.mark.parametrize("kernel,centering", [("IB_4", ia.SIDE), ("IB_6", ia.SIDE), ("BSPLINE_3", ia.CELL)])
def test_fp64_bitwise_with_every_shape(orc, kernel, centering):
    """Every FP64 tile shape (footprint and height) against the oracle, bit for bit."""
    from test_gpu_parity import _oracle_spread
    plan, X, F, shapes, dt = _setup(ia.FP64, kernel, centering)
    assert plan.tile_shape == FP64_SHAPES[0]
    level = ia.Level.uniform_periodic(40, 3)
    Xh, Fh = X.cpu().numpy(), F.cpu().numpy()
    want = _oracle_spread(orc, level, plan, [np.zeros(s) for s in shapes], Xh, Fh, kernel)
    for sh in FP64_SHAPES:
        plan.tile_shape = sh
        got = _spread(plan, X, F, shapes, dt)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), sh
    plan.tile_height = 16
    assert plan.tile_shape == (8, 8, 16)


FP32C_SHAPES = [(8, 8, 16), (8, 8, 8), (8, 8, 24), (16, 4, 16), (8, 8, 64), (8, 8, 128), (8, 8, 256)]


def _rel(a, b):
    return max(float(np.abs(x.astype(np.float64) - y).max()) for x, y in zip(a, b)) / \
        max(float(np.abs(y).max()) for y in b)


@pytest.mark.parametrize("kernel,centering", [("IB_6", ia.SIDE), ("IB_4", ia.SIDE), ("BSPLINE_3", ia.CELL)])
def test_fp32c_every_shape_and_autotune(kernel, centering):
    plan, X, F, shapes, dt = _setup(ia.FP32C, kernel, centering)
    assert plan.tile_shape == FP32C_SHAPES[0]
    ref = _spread(plan, X, F, shapes, dt)
    assert max(np.max(np.abs(a)) for a in ref) > 0
    # the FP64 spread of the same points: the FP32C bar
    p64, _, _, _, _ = _setup(ia.FP64, kernel, centering)
    f64 = _spread(p64, X, F.double(), shapes, torch.float64)
    for sh in FP32C_SHAPES:
        plan.tile_shape = sh
        got = _spread(plan, X, F, shapes, dt)
        again = _spread(plan, X, F, shapes, dt)
        for a, b in zip(got, again):
            assert np.array_equal(a, b), sh   # reproducible for a given shape
        assert _rel(got, f64) <= 1e-6, sh
    f = [torch.zeros(s, dtype=dt, device=DEV) for s in shapes]
    plan.bind(0, f)
    times = plan.autotune(0, F, X, reps=2)
    print("autotune", times, flush=True)
    assert sorted(times) == sorted(FP32C_SHAPES)
    assert all(t > 0 for t in times.values())
    assert plan.tile_shape == min(times, key=times.get)


@pytest.mark.parametrize("kernel", ["IB_6", "BSPLINE_3"])
def test_fp32c_point_order(kernel):
    """The points handed over in another order: the same grid within the FP32C
    tolerance (the chunks group other points), and the same bits again for a
    repeat of that order."""
    plan, X, F, shapes, dt = _setup(ia.FP32C, kernel, ia.SIDE, seed=9)
    ref = _spread(plan, X, F, shapes, dt)
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    perm = torch.randperm(X.shape[0], generator=g, device=DEV)
    Xp, Fp = X[perm].contiguous(), F[perm].contiguous()
    plan.bin(Xp)
    got = _spread(plan, Xp, Fp, shapes, dt)
    assert _rel(got, [a.astype(np.float64) for a in ref]) <= 1e-6
    plan.bin(Xp)
    again = _spread(plan, Xp, Fp, shapes, dt)
    for a, b in zip(got, again):
        assert np.array_equal(a, b)


def test_shape_validation():
    plan, X, F, shapes, dt = _setup(ia.FP64, "IB_4", ia.SIDE, n=16, npts=100)
    for bad in (0, 8, 12, 64):
        with pytest.raises(ia.IBAMRError):
            plan.tile_height = bad
    for bad in ((16, 16, 64), (8, 8, 8), (8, 16, 32)):
        with pytest.raises(ia.IBAMRError):
            plan.tile_shape = bad
    p32 = ia.LevelPlan(ia.Level.uniform_periodic(16, 3), "IB_4", ia.SIDE, ia.FP32C, max_points=10)
    for bad in ((8, 8, 32), (16, 16, 16)):
        with pytest.raises(ia.IBAMRError):
            p32.tile_shape = bad
    assert plan.tile_shape == (8, 8, 32)
    level = ia.Level.uniform_periodic(32, 2)
    p2 = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP32C, max_points=10)
    assert p2.tile_shape == (16, 16, 1)
    with pytest.raises(ia.IBAMRError):
        p2.tile_height = 16
    p3 = ia.LevelPlan(ia.Level.uniform_periodic(16, 3), "PIECEWISE_LINEAR", ia.SIDE, ia.FP32C, max_points=10)
    assert p3.tile_shape == (8, 8, 16)


def test_graph_keeps_its_shape_and_graph_after_autotune():
    """A graph captured with one tile shape replays it after another shape is
    chosen (the lists live as long as the plan); a graph captured after the
    autotuner replays the chosen one; both equal the eager calls on every
    replay."""
    plan, X, F, shapes, dt = _setup(ia.FP32C, "IB_4", ia.SIDE)
    f = [torch.zeros(s, dtype=dt, device=DEV) for s in shapes]
    u = [torch.rand(s, dtype=torch.float64, device=DEV).to(dt) for s in shapes]
    plan.bind(0, f)
    plan.bind(1, u)
    U = torch.zeros((X.shape[0], 3), dtype=dt, device=DEV)
    plan.spread(0, F, X, mode=ia.OVERWRITE)
    plan.interp(1, U, None)
    torch.cuda.synchronize()
    f_ref, U_ref = [a.clone() for a in f], U.clone()

    def capture():
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                plan.spread(0, F, X, mode=ia.OVERWRITE, stream=side)
                plan.interp(1, U, None, stream=side)
        torch.cuda.current_stream().wait_stream(side)
        return g

    def check(g, f_expect):
        for rep in range(3):
            for a in f:
                a.fill_(7.0)
            U.fill_(7.0)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for a, b in zip(f, f_expect):
                assert torch.equal(a, b), rep
            assert torch.equal(U, U_ref), rep
            plan.spread(0, F, X, mode=ia.OVERWRITE)   # eager work between replays
            torch.cuda.synchronize()

    g1 = capture()   # the default shape: f_ref
    plan.tile_shape = (8, 8, 8)
    check(g1, f_ref)
    plan.autotune(0, F, X, reps=1)
    plan.spread(0, F, X, mode=ia.OVERWRITE)   # the chosen shape, eagerly
    torch.cuda.synchronize()
    f_tuned = [a.clone() for a in f]
    g2 = capture()
    check(g2, f_tuned)
    check(g1, f_ref)
