"""GPU physical boundary operators (ibamr_plan_accumulate_physical_boundary /
ibamr_plan_fill_physical_boundary) against the oracle's restatement of the
LINEAR Robin operators (bitwise in FP64, per patch), their validation of the
boxes, and the spread near walls: with Neumann conditions the fold keeps the
whole spread force inside the domain."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from test_gpu_parity import _dev  # noqa: E402

pytestmark = pytest.mark.gpu

KINDS = {"dirichlet": (1.0, 0.0, 0.0), "neumann": (0.0, 1.0, 0.0)}


def _level(ndim, n, cuts=None, periodic=()):
    h = 1.0 / n
    cuts = cuts or [[]] * ndim
    import itertools
    edges = [[0] + list(c) + [n] for c in cuts]
    patches = []
    for idx in itertools.product(*[range(len(e) - 1) for e in edges]):
        lo = [edges[d][idx[d]] for d in range(ndim)]
        hi = [edges[d][idx[d] + 1] - 1 for d in range(ndim)]
        patches.append(ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]))
    return ia.Level([h] * ndim, [0] * ndim, [n - 1] * ndim, [0.0] * ndim, [1.0] * ndim,
                    [n if d in periodic else 0 for d in range(ndim)], patches)


def _boxes_and_coefs(level, gcw, narr, rng):
    boxes, abg = [], []
    for p in range(len(level.patches)):
        for (c, l, lo, hi) in ia.physical_boundary_boxes(level, p, gcw):
            boxes.append((p, c, l, lo, hi))
            row = []
            for _ in range(narr):
                k = rng.integers(0, 3)
                row.append(KINDS["dirichlet"] if k == 0 else KINDS["neumann"] if k == 1 else
                           (rng.uniform(0.5, 2.0), rng.uniform(0.1, 1.0), rng.uniform(-1, 1)))
            abg.append(row)
    return boxes, np.array(abg).reshape(len(boxes), narr, 3)


@pytest.mark.parametrize("ndim,side,cuts", [(2, False, None), (3, False, None), (2, True, None), (3, True, None),
                                            (3, True, [[5], [4], []]), (3, False, [[], [6], [3]])])
@pytest.mark.parametrize("adjoint", [True, False])
def test_matches_oracle_fp64(orc, ndim, side, cuts, adjoint):
    print("phys bdry", ndim, side, cuts, adjoint, flush=True)
    n, gcw = 10, 3
    level = _level(ndim, n, cuts)
    depth = 1 if side else 2
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE if side else ia.CELL, ia.FP64, depth=depth, data_gcw=[gcw] * ndim,
                        max_points=1)
    rng = np.random.default_rng(ndim * 100 + side * 10 + adjoint)
    narr = ndim if side else depth
    boxes, abg = _boxes_and_coefs(level, gcw, narr, rng)
    plan.set_physical_boundary(boxes, abg)
    arrays = []
    for p in range(len(level.patches)):
        for a in range(ndim if side else 1):
            arrays.append(rng.standard_normal(ia.array_shape(level, p, [gcw] * ndim, plan.centering, axis=a,
                                                             depth=depth)))
    dev = [_dev(a) for a in arrays]
    plan.bind(0, dev)
    if adjoint:
        plan.accumulate_physical_boundary(0)
    else:
        plan.fill_physical_boundary(0)
    torch.cuda.synchronize()
    na = ndim if side else 1
    for p, P in enumerate(level.patches):
        own = [k for k, b in enumerate(boxes) if b[0] == p]
        ref = [arrays[p * na + a].copy() for a in range(na)]
        orc.phys_bdry(ref, P.ilower, P.iupper, gcw, level.dx, [boxes[k][1:] for k in own], abg[own], side, adjoint,
                      depth=depth)
        for a in range(na):
            assert np.array_equal(dev[p * na + a].cpu().numpy(), ref[a]), (p, a)
    plan.close()


def test_fp32c_close_to_fp64(orc):
    n, gcw, ndim = 12, 3, 3
    level = _level(ndim, n)
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP32C, data_gcw=[gcw] * 3, max_points=1)
    rng = np.random.default_rng(9)
    boxes, abg = _boxes_and_coefs(level, gcw, 3, rng)
    plan.set_physical_boundary(boxes, abg)
    arrays = [rng.standard_normal(ia.array_shape(level, 0, [gcw] * 3, ia.SIDE, axis=a)) for a in range(3)]
    dev = [_dev(a.astype(np.float32)) for a in arrays]
    plan.bind(0, dev)
    plan.accumulate_physical_boundary(0)
    torch.cuda.synchronize()
    ref = [a.astype(np.float32).astype(np.float64) for a in arrays]
    orc.phys_bdry(ref, [0] * 3, [n - 1] * 3, gcw, level.dx, [b[1:] for b in boxes], abg, True, True)
    for a in range(3):
        got = dev[a].cpu().numpy().astype(np.float64)
        assert np.max(np.abs(got - ref[a])) <= 1e-6 * np.max(np.abs(ref[a]))
    plan.close()


def test_box_validation():
    level = _level(3, 8)
    plan = ia.LevelPlan(level, "IB_4", ia.CELL, ia.FP64, data_gcw=[3] * 3, max_points=1)
    one = np.zeros((1, 1, 3))
    with pytest.raises(ia.IBAMRError, match="not inside the ghost box"):
        plan.set_physical_boundary([(0, 2, 0, [0, -4, -3], [7, -1, -1])], one)
    with pytest.raises(ia.IBAMRError, match="cell layer just outside its face"):
        plan.set_physical_boundary([(0, 1, 0, [0, 0, 0], [0, 7, 7])], one)
    with pytest.raises(ia.IBAMRError, match="outside the patch on its sides"):
        plan.set_physical_boundary([(0, 3, 7, [5, 5, 5], [10, 10, 10])], one)
    with pytest.raises(ia.IBAMRError, match="bad location"):
        plan.set_physical_boundary([(0, 1, 6, [-1, 0, 0], [-1, 7, 7])], one)
    plan.close()


def test_spread_near_walls_neumann_fold_keeps_force(orc):
    """Cell-centred spread in a walled box (one patch), then the fold with
    homogeneous Neumann conditions: the interior sum equals sum F / h^3."""
    n, ndim = 24, 3
    h = 1.0 / n
    level = _level(ndim, n)
    gw = ia.get_minimum_ghost_width("BSPLINE_3")
    plan = ia.LevelPlan(level, "BSPLINE_3", ia.CELL, ia.FP64, depth=3, data_gcw=[gw] * 3, max_points=4000)
    rng = np.random.default_rng(12)
    X = rng.uniform(0.0, 1.0, (4000, 3))
    X[:1000] = rng.uniform(0.0, 1.5 * h, (1000, 3))          # crowd the corner at the origin
    F = rng.standard_normal((4000, 3))
    f = [torch.zeros(ia.array_shape(level, 0, [gw] * 3, ia.CELL, depth=3), dtype=torch.float64, device="cuda")]
    plan.bind(0, f)
    Xd = _dev(X)
    plan.bin(Xd)
    plan.spread(0, _dev(F), Xd, mode=ia.OVERWRITE)
    boxes = [(0,) + b for b in ia.physical_boundary_boxes(level, 0, gw)]
    plan.set_physical_boundary(boxes, np.tile([0.0, 1.0, 0.0], (len(boxes), 3, 1)))
    plan.accumulate_physical_boundary(0)
    torch.cuda.synchronize()
    g = f[0].cpu().numpy()
    inner = g[:, gw:gw + n, gw:gw + n, gw:gw + n]
    total = inner.reshape(3, -1).sum(axis=1) * h ** 3
    assert np.allclose(total, F.sum(axis=0), rtol=1e-12, atol=1e-12)
    plan.close()


def _varying_coefs(rng, boxes, narr, ndim, side):
    """{box index: [(3, n) array per array j]}: every column of a codim-1 box
    its own Dirichlet, Neumann or Robin coefficients and its own g."""
    out = {}
    for k, b in enumerate(boxes):
        if b[1] != 1:
            continue
        per = []
        for j in range(narr):
            m = orc_ncoef(ndim, b[1:], j, side)
            c = np.zeros((3, m))
            kind = rng.integers(0, 3, m)
            c[0] = np.where(kind == 0, 1.0, np.where(kind == 1, 0.0, rng.uniform(0.5, 2.0, m)))
            c[1] = np.where(kind == 0, 0.0, np.where(kind == 1, 1.0, rng.uniform(0.1, 1.0, m)))
            c[2] = rng.uniform(-1.0, 1.0, m)
            per.append(c)
        out[k] = per
    return out


def orc_ncoef(ndim, box, j, side):
    import oracle
    return oracle.phys_bdry_ncoef(ndim, box, j, side)


@pytest.mark.parametrize("ndim,side,cuts", [(2, False, None), (3, False, None), (2, True, None), (3, True, None),
                                            (3, True, [[5], [4], []]), (2, False, [[4], [6]])])
@pytest.mark.parametrize("adjoint", [True, False])
def test_varying_coefficients_match_oracle_fp64(orc, ndim, side, cuts, adjoint):
    """Coefficients varying over each face (RobinBcCoefStrategy::setBcCoefs
    filling the coefficient box): bitwise against the oracle, patch by patch."""
    print("phys bdry varying", ndim, side, cuts, adjoint, flush=True)
    n, gcw = 10, 3
    level = _level(ndim, n, cuts)
    depth = 1 if side else 2
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE if side else ia.CELL, ia.FP64, depth=depth, data_gcw=[gcw] * ndim,
                        max_points=1)
    rng = np.random.default_rng(500 + ndim * 100 + side * 10 + adjoint)
    narr = ndim if side else depth
    boxes = [(p,) + b for p in range(len(level.patches)) for b in ia.physical_boundary_boxes(level, p, gcw)]
    cf = _varying_coefs(rng, boxes, narr, ndim, side)
    plan.set_physical_boundary_coefs(boxes, [c for k in sorted(cf) for c in cf[k]])
    na = ndim if side else 1
    arrays = [rng.standard_normal(ia.array_shape(level, p, [gcw] * ndim, plan.centering, axis=a, depth=depth))
              for p in range(len(level.patches)) for a in range(na)]
    dev = [_dev(a) for a in arrays]
    plan.bind(0, dev)
    if adjoint:
        plan.accumulate_physical_boundary(0)
    else:
        plan.fill_physical_boundary(0)
    torch.cuda.synchronize()
    for p, P in enumerate(level.patches):
        own = [k for k, b in enumerate(boxes) if b[0] == p]
        pc = [c for k in own if k in cf for c in cf[k]]
        ref = [arrays[p * na + a].copy() for a in range(na)]
        orc.phys_bdry_coefs(ref, P.ilower, P.iupper, gcw, level.dx, [boxes[k][1:] for k in own],
                            np.concatenate([c.reshape(-1) for c in pc]) if pc else np.zeros(0), side, adjoint,
                            depth=depth)
        for a in range(na):
            assert np.array_equal(dev[p * na + a].cpu().numpy(), ref[a]), (p, a)
    plan.close()


@pytest.mark.parametrize("side", [False, True])
def test_constant_coefficient_arrays_equal_uniform(side):
    """Arrays holding one (a, b, g) per box and array act exactly as the uniform
    entry point with those values."""
    ndim, n, gcw = 3, 9, 3
    level = _level(ndim, n, [[4], [], []])
    depth = 1 if side else 2
    narr = ndim if side else depth
    rng = np.random.default_rng(77 + side)
    boxes = [(p,) + b for p in range(len(level.patches)) for b in ia.physical_boundary_boxes(level, p, gcw)]
    _, abg = _boxes_and_coefs(level, gcw, narr, rng)
    cf = []
    for k, b in enumerate(boxes):
        if b[1] == 1:
            cf += [np.repeat(abg[k, j][:, None], orc_ncoef(ndim, b[1:], j, side), axis=1) for j in range(narr)]
    na = ndim if side else 1
    arrays = [rng.standard_normal(ia.array_shape(level, p, [gcw] * ndim, ia.SIDE if side else ia.CELL, axis=a,
                                                 depth=depth))
              for p in range(len(level.patches)) for a in range(na)]
    out = []
    for varying in (False, True):
        plan = ia.LevelPlan(level, "IB_4", ia.SIDE if side else ia.CELL, ia.FP64, depth=depth,
                            data_gcw=[gcw] * ndim, max_points=1)
        if varying:
            plan.set_physical_boundary_coefs(boxes, cf)
        else:
            plan.set_physical_boundary(boxes, abg)
        dev = [_dev(a) for a in arrays]
        plan.bind(0, dev)
        plan.accumulate_physical_boundary(0)
        plan.fill_physical_boundary(0)
        torch.cuda.synchronize()
        out.append([d.cpu().numpy() for d in dev])
        plan.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_coefficient_count_checked():
    level = _level(3, 8)
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, data_gcw=[3] * 3, max_points=1)
    boxes = [(0,) + b for b in ia.physical_boundary_boxes(level, 0, 3)]
    need = sum(3 * orc_ncoef(3, b[1:], j, True) for b in boxes if b[1] == 1 for j in range(3))
    plan.set_physical_boundary_coefs(boxes, [np.zeros(need)])
    with pytest.raises(ia.IBAMRError, match="values given"):
        plan.set_physical_boundary_coefs(boxes, [np.zeros(need - 1)])
    plan.close()
