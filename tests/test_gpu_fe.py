"""GPU FE point source (ibamr_fe_quadrature_points) against the oracle's
restatement of FEDataManager::spread's quadrature loop (bitwise in FP64), and
the FE spread end to end: quadrature points through a level plan equal the
oracle's spread of the same points, and the spread force sums to the integral
of the force density."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from ibamr_amd import fe  # noqa: E402

pytestmark = pytest.mark.gpu


def _mesh(ndim, n, seed):
    lo, hi = [0.3, 0.35, 0.25][:ndim], [0.62, 0.6, 0.7][:ndim]
    X, E = fe.box_mesh(ndim, n, lo, hi)
    rng = np.random.default_rng(seed)
    X = X + rng.uniform(-0.2, 0.2, X.shape) * (np.array(hi) - np.array(lo)) / np.array(n)  # distorted elements
    return X, E


@pytest.mark.parametrize("ndim,n,g", [(2, (40, 30), 3), (3, (20, 16, 24), 2)])
def test_quadrature_points_bitwise(orc, ndim, n, g):
    print("fe qp", ndim, n, g, flush=True)
    X, E = _mesh(ndim, n, 11)
    xi, w = fe.gauss_rule(ndim, g)
    phi = fe.q1_shape(ndim, xi)
    J = fe.q1_jxw(X, E, xi, w)
    F = np.random.default_rng(12).standard_normal((len(X), ndim))
    Xq, Fq = orc.fe_quadrature_points(E, phi, J, X, F)
    c = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
    gX, gF = ia.fe_quadrature_points(c(E, torch.int32), c(phi), c(J), c(X), c(F))
    assert np.array_equal(gX.cpu().numpy(), Xq) and np.array_equal(gF.cpu().numpy(), Fq)
    gX2, gF2 = ia.fe_quadrature_points(c(E, torch.int32), c(phi), c(J), c(X), c(F), f_dtype=torch.float32)
    assert np.array_equal(gX2.cpu().numpy(), Xq) and np.array_equal(gF2.cpu().numpy(), Fq.astype(np.float32))
    gX3, none = ia.fe_quadrature_points(c(E, torch.int32), c(phi), None, c(X))
    assert none is None and np.array_equal(gX3.cpu().numpy(), Xq)


def test_fe_spread_end_to_end(orc):
    N = 48
    h = 1.0 / N
    X, E = _mesh(3, (12, 10, 14), 13)
    xi, w = fe.gauss_rule(3, 2)
    phi = fe.q1_shape(3, xi)
    J = fe.q1_jxw(X, E, xi, w)
    F = np.random.default_rng(14).standard_normal((len(X), 3))
    c = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
    Xq, Fq = ia.fe_quadrature_points(c(E, torch.int32), c(phi), c(J), c(X), c(F))
    level = ia.Level.uniform_periodic(N, 3)
    gw = ia.get_minimum_ghost_width("IB_4")
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=Xq.shape[0])
    f = [torch.zeros(ia.array_shape(level, 0, [gw] * 3, ia.SIDE, axis=a), dtype=torch.float64, device="cuda")
         for a in range(3)]
    plan.bind(0, f)
    plan.bin(Xq)
    plan.spread(0, Fq, Xq, mode=ia.OVERWRITE)
    torch.cuda.synchronize()
    Xn, Fn = Xq.cpu().numpy(), Fq.cpu().numpy()
    il, iu = [0] * 3, [N - 1] * 3
    ref = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    cells = orc.get_cell_index(Xn, [0.0] * 3, [1.0] * 3, [h] * 3, il, iu)
    idx, sh = orc.build_patch_entries(cells, il, iu, gw, [N] * 3, il, [h] * 3, [8] * 3)
    orc.spread_side("IB_4", [h] * 3, [0.0] * 3, il, iu, [gw] * 3, ref, Xn, Fn, indices=idx, Xshift=sh)
    total = Fn.sum(axis=0)
    for a in range(3):
        got = f[a].cpu().numpy()
        sl = tuple(slice(gw, gw + N + (1 if (2 - ax) == a else 0)) for ax in range(3))
        assert np.max(np.abs(got[sl] - ref[a][sl])) <= 1e-13 * np.max(np.abs(ref[a][sl]))
        # the periodic side array holds face N == face 0: count the faces 0..N-1
        inner = tuple(slice(gw, gw + N) for _ in range(3))
        assert abs(got[inner].sum() * h ** 3 - total[a]) <= 1e-12 * np.abs(Fn[:, a]).sum()
    plan.close()


@pytest.mark.parametrize("ndim,n", [(2, (40, 30)), (3, (20, 16, 24))])
def test_interp_rhs_bitwise(orc, ndim, n):
    print("fe rhs", ndim, n, flush=True)
    X, E = _mesh(ndim, n, 21)
    xi, w = fe.gauss_rule(ndim, 2)
    phi = fe.q1_shape(ndim, xi)
    J = fe.q1_jxw(X, E, xi, w)
    rng = np.random.default_rng(22)
    U = rng.standard_normal((len(E) * len(w), ndim))
    F0 = rng.standard_normal((len(X), ndim))
    c = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
    mesh = ia.FEMesh(E, len(X))
    Fd = torch.full((len(X), ndim), float("nan"), dtype=torch.float64, device="cuda")
    mesh.integrate_rhs(Fd, c(U), c(phi), c(J))
    assert np.array_equal(Fd.cpu().numpy(), orc.fe_integrate_rhs(E, phi, J, U, len(X)))
    Fd = c(F0)
    mesh.integrate_rhs(Fd, c(U, torch.float32), c(phi), c(J), accumulate=True)
    want = orc.fe_integrate_rhs(E, phi, J, U.astype(np.float32).astype(np.float64), len(X), F=F0)
    assert np.array_equal(Fd.cpu().numpy(), want)
    mesh.close()


def test_fe_mesh_rejects_bad_node():
    with pytest.raises(ia.IBAMRError, match="element 1: bad node index"):
        ia.FEMesh(np.array([[0, 1, 2, 3], [1, 2, 3, 9]]), 5)


def test_fe_interp_end_to_end(orc):
    """Velocity at the quadrature points by the level plan, then the weighted
    right-hand side: equals the oracle's interpolation + integration."""
    N = 48
    h = 1.0 / N
    X, E = _mesh(3, (10, 12, 9), 23)
    xi, w = fe.gauss_rule(3, 2)
    phi = fe.q1_shape(3, xi)
    J = fe.q1_jxw(X, E, xi, w)
    c = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
    Xq, _ = ia.fe_quadrature_points(c(E, torch.int32), c(phi), None, c(X))
    level = ia.Level.uniform_periodic(N, 3)
    gw = ia.get_minimum_ghost_width("BSPLINE_3")
    plan = ia.LevelPlan(level, "BSPLINE_3", ia.SIDE, ia.FP64, max_points=Xq.shape[0])
    g = torch.Generator(device="cuda")
    g.manual_seed(24)
    u = [torch.rand(ia.array_shape(level, 0, [gw] * 3, ia.SIDE, axis=a), generator=g, dtype=torch.float64,
                    device="cuda") for a in range(3)]
    plan.bind(1, u)
    plan.fill_periodic_ghosts(1)
    plan.bin(Xq)
    Uq = torch.empty_like(Xq)
    plan.interp(1, Uq, Xq)
    mesh = ia.FEMesh(E, len(X))
    Fd = torch.empty((len(X), 3), dtype=torch.float64, device="cuda")
    mesh.integrate_rhs(Fd, Uq, c(phi), c(J))
    torch.cuda.synchronize()
    Xn = Xq.cpu().numpy()
    un = [a.cpu().numpy() for a in u]
    il, iu = [0] * 3, [N - 1] * 3
    Uref = orc.interp_side("BSPLINE_3", [h] * 3, [0.0] * 3, il, iu, [gw] * 3, un, Xn)
    assert np.max(np.abs(Uq.cpu().numpy() - Uref)) <= 1e-14
    want = orc.fe_integrate_rhs(E, phi, J, Uref, len(X))
    assert np.max(np.abs(Fd.cpu().numpy() - want)) <= 1e-13 * np.abs(want).max()
    mesh.close()
    plan.close()


def _c(a, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()


@pytest.mark.parametrize("ndim,n", [(2, (40, 30)), (3, (20, 16, 24))])
def test_nodal_quadrature_mass_and_weights_bitwise(orc, ndim, n):
    """Nodal quadrature: the diagonal mass of the nodal rule (and of a Gauss rule)
    bitwise against the oracle; F (x) dX and U * dX into a right-hand side."""
    print("fe nodal", ndim, n, flush=True)
    X, E = _mesh(ndim, n, 21)
    m = ia.FEMesh(E, len(X))
    for xi, w in (fe.nodal_rule(ndim), fe.gauss_rule(ndim, 2)):
        phi = fe.q1_shape(ndim, xi)
        J = fe.q1_jxw(X, E, xi, w)
        M = m.nodal_mass(_c(phi), _c(J))
        want = orc.fe_nodal_mass(E, phi, J, len(X))
        assert np.array_equal(M.cpu().numpy(), want)
    rng = np.random.default_rng(22)
    F = rng.standard_normal((len(X), ndim))
    FdX = ia.fe_nodal_weight(_c(F), M)
    assert np.array_equal(FdX.cpu().numpy(), F * want[:, None])
    F32 = ia.fe_nodal_weight(_c(F), M, out=torch.empty(F.shape, dtype=torch.float32, device="cuda"))
    assert np.array_equal(F32.cpu().numpy(), (F * want[:, None]).astype(np.float32))
    R0 = rng.standard_normal((len(X), ndim))
    R = _c(R0)
    ia.fe_nodal_weight(_c(F.astype(np.float32), torch.float32), M, out=R, accumulate=True)
    assert np.array_equal(R.cpu().numpy(), R0 + F.astype(np.float32).astype(np.float64) * want[:, None])
    m.close()


def test_nodal_quadrature_spread_keeps_the_force(orc):
    """Nodes as points with F (x) dX through a periodic level plan: the spread
    force integrates to sum_n F_n dX_n."""
    N = 48
    h = 1.0 / N
    X, E = _mesh(3, (12, 10, 14), 23)
    xi, w = fe.nodal_rule(3)
    m = ia.FEMesh(E, len(X))
    M = m.nodal_mass(_c(fe.q1_shape(3, xi)), _c(fe.q1_jxw(X, E, xi, w)))
    F = np.random.default_rng(24).standard_normal((len(X), 3))
    P = ia.fe_nodal_weight(_c(F), M)
    level = ia.Level.uniform_periodic(N, 3)
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, max_points=len(X))
    f = [torch.zeros(ia.array_shape(level, 0, plan.data_gcw, ia.SIDE, axis=a), dtype=torch.float64, device="cuda")
         for a in range(3)]
    plan.bind(0, f)
    Xd = _c(X)
    plan.bin(Xd)
    plan.spread(0, P, Xd, mode=ia.OVERWRITE)
    torch.cuda.synchronize()
    g = plan.data_gcw[0]
    FM = F * M.cpu().numpy()[:, None]
    for a in range(3):   # the periodic side array holds face N == face 0: count the faces 0..N-1
        tot = float(f[a][g:g + N, g:g + N, g:g + N].sum()) * h ** 3
        assert abs(tot - FM[:, a].sum()) <= 1e-12 * np.abs(FM[:, a]).sum()
    m.close()


TYPES = {"EDGE2": 2, "EDGE3": 3, "TRI3": 3, "TRI6": 6, "QUAD4": 4, "QUAD8": 8, "QUAD9": 9, "TET4": 4, "TET10": 10,
         "HEX8": 8, "HEX20": 20, "HEX27": 27}


@pytest.mark.parametrize("etype", sorted(TYPES))
def test_adaptive_orders_bitwise(orc, etype):
    """getQuadratureKey's adaptive order of 20000 random elements of each type
    (node coordinates spread over several cell widths, both rule families,
    including densities where npts falls below its floor and above the QGAUSS
    cap)."""
    rng = np.random.default_rng(31 + len(etype))
    nb = TYPES[etype]
    ndim = 3
    X = rng.uniform(0.0, 1.0, (5000, ndim)) * rng.choice([0.01, 0.1, 1.0, 10.0], (5000, 1))
    E = rng.integers(0, len(X), (20000, nb)).astype(np.int32)
    for qt in (ia.FE_QGAUSS, ia.FE_QGRID):
        for dens in (0.3, 2.0):
            got = ia.fe_adaptive_quadrature_orders(etype, _c(E, torch.int32), _c(X), qt, dens, 0.05)
            want = orc.fe_adaptive_orders(etype, E, X, qt, dens, 0.05)
            assert np.array_equal(got.cpu().numpy(), want), (qt, dens)
    with pytest.raises(ia.IBAMRError, match="only for QGAUSS or QGRID"):
        ia.fe_adaptive_quadrature_orders(etype, _c(E, torch.int32), _c(X), 5, 1.0, 0.05)


@pytest.mark.parametrize("ndim,n", [(2, (30, 20)), (3, (12, 10, 8))])
def test_rules_per_element_bitwise(orc, ndim, n):
    """Adaptive orders of a distorted mesh pick a Gauss rule per element; the
    quadrature points, F JxW and the weighted right-hand side with those rules
    are bitwise the oracle's."""
    print("fe rules", ndim, n, flush=True)
    X, E = _mesh(ndim, n, 41)
    etype = "QUAD4" if ndim == 2 else "HEX8"
    h = (0.3 / np.array(n)).min()
    orders = ia.fe_adaptive_quadrature_orders(etype, _c(E, torch.int32), _c(X), ia.FE_QGAUSS, 2.0, h).cpu().numpy()
    assert np.array_equal(orders, orc.fe_adaptive_orders(etype, E, X, 0, 2.0, h))
    uniq = sorted(set(orders.tolist()))
    assert len(uniq) >= 2, uniq              # the distortion gives several rules
    rules = [fe.gauss_rule_for_order(ndim, o) for o in uniq]
    phis = [fe.q1_shape(ndim, xi) for xi, _ in rules]
    er = np.array([uniq.index(o) for o in orders], dtype=np.int32)
    J = np.concatenate([fe.q1_jxw(X, E[e:e + 1], *rules[r]).reshape(-1) for e, r in enumerate(er)])
    F = np.random.default_rng(42).standard_normal((len(X), 2))
    R = ia.FERules(phis, er)
    gX, gF = ia.fe_quadrature_points_rules(_c(E, torch.int32), R, _c(J), _c(X), _c(F))
    Xq, Fq = orc.fe_quadrature_points_rules(E, phis, er, J, X, F)
    assert np.array_equal(gX.cpu().numpy(), Xq) and np.array_equal(gF.cpu().numpy(), Fq)
    U = np.random.default_rng(43).standard_normal((R.nqp_total, 2))
    m = ia.FEMesh(E, len(X))
    G = torch.zeros((len(X), 2), dtype=torch.float64, device="cuda")
    m.integrate_rhs_rules(G, _c(U), R, _c(J))
    assert np.array_equal(G.cpu().numpy(), orc.fe_integrate_rhs_rules(E, phis, er, J, U, len(X)))
    m.close()
