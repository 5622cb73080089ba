"""GPU redistribution (ibamr_plan_redistribute + ibamr_permute_rows) against
the oracle's restatement of LDataManager::begin/endDataRedistribution: moved
positions bitwise, periodic offsets, the new order and the per-patch offsets
exactly; error reports; and a spread after the permutation equals the spread
before it (up to summation order)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from ibamr_amd import synthetic  # noqa: E402

pytestmark = pytest.mark.gpu


def _tiled_level(ndim, N, splits, periodic=True):
    """Patches tiling [0, N)^ndim: axis d cut at splits[d]."""
    h = 1.0 / N
    edges = [[0] + list(s) + [N] for s in splits]
    import itertools
    patches = []
    for idx in itertools.product(*[range(len(e) - 1) for e in edges]):
        lo = [edges[d][idx[d]] for d in range(ndim)]
        hi = [edges[d][idx[d] + 1] - 1 for d in range(ndim)]
        patches.append(ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]))
    return ia.Level([h] * ndim, [0] * ndim, [N - 1] * ndim, [0.0] * ndim, [1.0] * ndim,
                    [N if periodic else 0] * ndim, patches)


def _oracle(orc, level, X, lag=None, error_if_leave=False):
    boxes = [(p.ilower, p.iupper) for p in level.patches]
    return orc.redistribute(X, level.domain_x_lower, level.domain_x_upper, level.dx, level.domain_lower,
                            level.domain_upper, [s > 0 for s in level.periodic_shift], boxes, lag=lag,
                            error_if_leave=error_if_leave)


@pytest.mark.parametrize("ndim,n,use_lag", [(2, 5000, True), (3, 200000, True), (3, 100000, False),
                                            (3, 0, False)])
def test_redistribute_matches_oracle(orc, ndim, n, use_lag):
    print("redistribute", ndim, n, use_lag, flush=True)
    N = 64
    level = _tiled_level(ndim, N, [[16, 40]] * (ndim - 1) + [[8, 24, 48]])
    rng = np.random.default_rng(100 + ndim + n)
    X = rng.uniform(-1.5, 2.5, (n, ndim))
    X[: n // 10] = rng.uniform(0.0, 1.0, (n // 10, ndim))
    if n > 10:
        X[1, 0] = 1.0
        X[2, 0] = -1e-17
        X[3:8] = X[8:13]  # coincident points
    lag = rng.permutation(n).astype(np.int32) + 3 if use_lag else None
    rc, Xw, order, poff, offs = _oracle(orc, level, X, lag)
    assert rc == 0
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=max(n, 1))
    Xd = torch.from_numpy(X.copy()).cuda()
    lagd = torch.from_numpy(lag).cuda() if use_lag else None
    got_order, got_poff, got_offs = plan.redistribute(Xd, lagd, periodic_offsets=True)
    torch.cuda.synchronize()
    assert np.array_equal(Xd.cpu().numpy(), Xw)
    assert np.array_equal(got_order.cpu().numpy(), order)
    assert np.array_equal(got_poff.cpu().numpy(), poff)
    assert np.array_equal(got_offs.cpu().numpy(), offs)
    # the LData arrays move with the order (VecScatter)
    if n:
        F = torch.randn((n, ndim), dtype=torch.float32, device="cuda")
        Xp = torch.empty_like(Xd)
        Fp = torch.empty_like(F)
        ia.permute_rows(Xp, Xd, got_order)
        ia.permute_rows(Fp, F, got_order)
        o = got_order.long()
        assert torch.equal(Xp, Xd[o]) and torch.equal(Fp, F[o])
    plan.close()


def test_redistribute_errors(orc):
    level = _tiled_level(3, 32, [[16]] * 3, periodic=False)
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=16)
    X = torch.tensor([[0.5, 0.5, 0.5], [1.2, 0.5, 0.5]], dtype=torch.float64, device="cuda")
    with pytest.raises(ia.IBAMRError, match="escaped"):
        plan.redistribute(X.clone(), error_if_points_leave=True)
    Xc = X.clone()
    plan.redistribute(Xc)  # clamped instead
    assert float(Xc[1, 0]) == 1.0 - np.sqrt(np.finfo(float).eps)
    rc, Xw, *_ = _oracle(orc, level, X.cpu().numpy())
    assert rc == 0 and np.array_equal(Xc.cpu().numpy(), Xw)
    # a fine level that does not cover the point
    sub = ia.Level(level.dx, level.domain_lower, level.domain_upper, level.domain_x_lower, level.domain_x_upper,
                   [0, 0, 0], level.patches[:4])
    p2 = ia.LevelPlan(sub, "IB_4", ia.SIDE, ia.FP64, max_points=16)
    with pytest.raises(ia.IBAMRError, match="in no patch"):
        p2.redistribute(torch.tensor([[0.9, 0.9, 0.9]], dtype=torch.float64, device="cuda"))
    pp = ia.LevelPlan(_tiled_level(3, 32, [[16]] * 3), "IB_4", ia.SIDE, ia.FP64, max_points=16)
    with pytest.raises(ia.IBAMRError, match="non-finite"):
        pp.redistribute(torch.tensor([[float("nan"), 0.5, 0.5]], dtype=torch.float64, device="cuda"))
    for p in (plan, p2, pp):
        p.close()


def test_redistribute_amr_level_then_spread(orc):
    """AMR fine level: redistribute, permute X and F, bin again; the spread
    equals the spread before redistribution to summation order."""
    level, X = synthetic.amr_level(coarse=32, ratio=2, npoints=1 << 16, nshell=3, seed=8, block=16)
    n = X.shape[0]
    rc, Xw, order, poff, _ = _oracle(orc, level, X)
    assert rc == 0
    gw = [ia.get_minimum_ghost_width("BSPLINE_3")] * 3
    plan = ia.LevelPlan(level, "BSPLINE_3", ia.SIDE, ia.FP64, data_gcw=gw, max_points=n)
    shapes = [[ia.array_shape(level, p, gw, ia.SIDE, axis=a) for a in range(3)] for p in range(len(level.patches))]
    f1 = [torch.zeros(s, dtype=torch.float64, device="cuda") for ps in shapes for s in ps]
    f2 = [torch.zeros_like(t) for t in f1]
    F = torch.randn((n, 3), dtype=torch.float64, device="cuda")
    Xd = torch.from_numpy(X.copy()).cuda()
    plan.bind(0, f1)
    plan.bind(1, f2)
    plan.bin(Xd)
    plan.spread(0, F, Xd)
    order_d, poff_d, _ = plan.redistribute(Xd)
    assert np.array_equal(order_d.cpu().numpy(), order) and np.array_equal(poff_d.cpu().numpy(), poff)
    Xp, Fp = torch.empty_like(Xd), torch.empty_like(F)
    ia.permute_rows(Xp, Xd, order_d)
    ia.permute_rows(Fp, F, order_d)
    plan.bin(Xp)
    plan.spread(1, Fp, Xp)
    torch.cuda.synchronize()
    for a, b in zip(f1, f2):
        scale = float(a.abs().max()) + 1e-300
        assert float((a - b).abs().max()) <= 1e-13 * scale
    # patch p's points are the contiguous block [poff[p], poff[p+1])
    cells = np.floor(Xp.cpu().numpy() / level.dx[0]).astype(int)
    for q, pt in enumerate(level.patches[:20]):
        blk = cells[poff[q]:poff[q + 1]]
        assert np.all((blk >= pt.ilower) & (blk <= pt.iupper))
    plan.close()
