"""Interpolation from static superbins: on a level whose patch owns whole
superbins, the 3-D interpolation stages each superbin's block from its cells
widened by the kernel's stencil start offsets instead of the points' exact
extents.  Positions on and one ulp either side of cell faces and cell centres
(where X / dx rounds to the neighbouring cell, so a stencil can start one cell
outside the block: such points are read from the array) must interpolate
bit for bit like the oracle, for every kernel, side- and cell-centred (FP64),
and within the FP32C tolerance of FP64 with the set-up pass's stencils."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from test_gpu_parity import _arrays, _dev, _oracle_interp, _periodic_fill_np  # noqa: E402

pytestmark = pytest.mark.gpu
N = 32
KERNELS = ["IB_4", "IB_6", "BSPLINE_3", "BSPLINE_4", "BSPLINE_5", "BSPLINE_6", "PIECEWISE_LINEAR"]


def _boundary_points(rng, n=N):
    """Each axis in turn on a face, a face -/+ 1 ulp, a centre -/+ 1 ulp (cells
    chosen at random, superbin edges included); the other axes random."""
    h = 1.0 / n
    cells = np.concatenate([rng.integers(0, n, 120), np.array([0, 15, 16, 31] * 6)])
    out = []
    for ax in range(3):
        for base in (cells * h, (cells + 0.5) * h):
            for v in (base, np.nextafter(base, -1.0), np.nextafter(base, 2.0)):
                X = rng.uniform(0.0, 1.0, (v.size, 3))
                X[:, ax] = np.mod(v, 1.0)
                out.append(X)
    X = np.concatenate(out)
    X[X >= 1.0] = 0.0
    return np.ascontiguousarray(X)


def _field(level, plan, rng):
    return [_periodic_fill_np(a, plan.data_gcw[0], N) for a in _arrays(level, plan, rng)]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("centering", [ia.SIDE, ia.CELL])
def test_static_superbins_fp64_bitwise(orc, kernel, centering):
    print("static superbins", kernel, centering, flush=True)
    rng = np.random.default_rng(91)
    level = ia.Level.uniform_periodic(N, 3)
    depth = 1 if centering == ia.SIDE else 2
    X = _boundary_points(rng)
    plan = ia.LevelPlan(level, kernel, centering, ia.FP64, depth=depth, max_points=X.shape[0])
    u = _field(level, plan, rng)
    ud = [_dev(a) for a in u]
    plan.bind(0, ud)
    Xd = _dev(X)
    plan.bin(Xd)
    fdepth = 3 if centering == ia.SIDE else depth
    U = torch.full((X.shape[0], fdepth), float("nan"), dtype=torch.float64, device="cuda")
    plan.interp(0, U, Xd)
    torch.cuda.synchronize()
    want = _oracle_interp(orc, level, plan, u, X, kernel)
    got = U.cpu().numpy()
    assert np.array_equal(got, want), np.nanmax(np.abs(got - want))


@pytest.mark.parametrize("kernel", ["IB_4", "IB_6"])
def test_static_superbins_fp32c_setup_stencils(orc, kernel):
    """FP32C after a spread: the interpolation reads the set-up pass's stencils
    (level frame) for the same boundary positions."""
    rng = np.random.default_rng(92)
    level = ia.Level.uniform_periodic(N, 3)
    X = _boundary_points(rng)
    n = X.shape[0]
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP32C, max_points=n)
    u = _field(level, plan, rng)
    u32 = [a.astype(np.float32) for a in u]
    f = [torch.zeros(a.shape, dtype=torch.float32, device="cuda") for a in u32]
    plan.bind(0, f)
    plan.bind(1, [_dev(a) for a in u32])
    Xd = _dev(X)
    plan.bin(Xd)
    F = _dev(rng.standard_normal((n, 3)).astype(np.float32))
    plan.spread(0, F, Xd, mode=ia.OVERWRITE)
    U = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    plan.interp(1, U, None)
    torch.cuda.synchronize()
    want = _oracle_interp(orc, level, plan, [a.astype(np.float64) for a in u32], X, kernel)
    err = np.abs(U.cpu().numpy().astype(np.float64) - want).max() / np.abs(want).max()
    assert err <= 1e-6, err
