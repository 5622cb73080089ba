"""The FE point source (quadrature points and F*JxW) pinned by what the
mathematics fixes: on an affine Q1 mesh the quadrature points are the mapped
Gauss points, a Q1-exact field integrates exactly (sum F*JxW = integral of F),
and the element volumes sum to the box volume."""
import numpy as np
import pytest

from ibamr_amd import fe


@pytest.mark.parametrize("ndim,n,g", [(2, (3, 4), 2), (3, (2, 3, 2), 2), (3, (2, 2, 2), 3)])
def test_quadrature_source_closed_forms(orc, ndim, n, g):
    lo, hi = np.array([0.1, 0.2, 0.3][:ndim]), np.array([0.7, 0.5, 0.9][:ndim])
    X, E = fe.box_mesh(ndim, n, lo, hi)
    xi, w = fe.gauss_rule(ndim, g)
    phi = fe.q1_shape(ndim, xi)
    assert np.allclose(phi.sum(axis=0), 1.0, atol=1e-15)
    J = fe.q1_jxw(X, E, xi, w)
    assert abs(J.sum() - np.prod(hi - lo)) <= 1e-14
    A = np.random.default_rng(3).standard_normal((ndim, 3))
    c = np.array([1.0, -2.0, 0.5])
    F = c + X @ A                          # linear: reproduced exactly by Q1
    Xq, Fq = orc.fe_quadrature_points(E, phi, J, X, F)
    # mapped Gauss points: element e's box [lo_e, hi_e], x = lo_e + (xi + 1)/2 (hi_e - lo_e)
    h = (hi - lo) / np.array(n)
    e0 = X[E[:, 0]]
    want = e0[:, None, :] + (xi[None, :, :] + 1.0) * 0.5 * h[None, None, :]
    assert np.allclose(Xq.reshape(len(E), -1, ndim), want, atol=1e-15)
    integral = np.prod(hi - lo) * (c + (0.5 * (lo + hi)) @ A)
    assert np.allclose(Fq.sum(axis=0), integral, rtol=1e-13, atol=1e-15)
    # positions-only pass (weights 1)
    Xq2, _ = orc.fe_quadrature_points(E, phi, None, X)
    assert np.array_equal(Xq2, Xq)


def test_distorted_element_jacobian():
    X, E = fe.box_mesh(3, (1, 1, 1), [0, 0, 0], [1, 1, 1])
    X[6] += [0.3, 0.2, 0.1]                # move the (1,1,1) corner: a trilinear element
    xi, w = fe.gauss_rule(3, 3)
    J = fe.q1_jxw(X, E, xi, w)
    # volume of the trilinear hexahedron by the divergence theorem on a fine sampling
    rng = np.random.default_rng(0)
    s = rng.uniform(-1, 1, (200000, 3))
    detJ = np.linalg.det(np.einsum("jkq,kd->qdj", fe.q1_shape_grad(3, s), X[E[0]]))
    assert abs(J.sum() - 8.0 * detJ.mean()) < 5e-3


@pytest.mark.parametrize("ndim,n", [(2, (5, 4)), (3, (3, 2, 4))])
def test_interp_rhs_is_adjoint_of_spread_source(orc, ndim, n):
    """<F, rhs(U)> = <F*JxW at the quadrature points, U>: the weighted
    interpolation is the transpose of the spread's point source; and U = c
    gives rhs summing to c * volume."""
    X, E = fe.box_mesh(ndim, n, [0.1] * ndim, [0.9] * ndim)
    X = X + np.random.default_rng(1).uniform(-0.02, 0.02, X.shape)
    xi, w = fe.gauss_rule(ndim, 2)
    phi = fe.q1_shape(ndim, xi)
    J = fe.q1_jxw(X, E, xi, w)
    rng = np.random.default_rng(2)
    F = rng.standard_normal((len(X), 2))
    U = rng.standard_normal((len(E) * len(w), 2))
    _, Fq = orc.fe_quadrature_points(E, phi, J, X, F)
    rhs = orc.fe_integrate_rhs(E, phi, J, U, len(X))
    assert abs((F * rhs).sum() - (Fq * U).sum()) <= 1e-13 * np.abs(F).sum() * np.abs(U).max()
    c = np.array([2.0, -0.5])
    rhs1 = orc.fe_integrate_rhs(E, phi, J, np.tile(c, (len(E) * len(w), 1)), len(X))
    assert np.allclose(rhs1.sum(axis=0), c * J.sum(), rtol=1e-13)


# ---- nodal quadrature: the diagonal mass of the nodal rule and of a Gauss rule

def _adjacent(E, nnodes):
    return np.bincount(E.ravel(), minlength=nnodes)


@pytest.mark.parametrize("ndim,n", [(2, (3, 4)), (3, (2, 3, 2))])
def test_nodal_mass_closed_forms(orc, ndim, n):
    """Nodal rule: M_i = (cell volume / 2^ndim) x (elements at node i), summing
    to the box volume; 2-point Gauss: M_i = prod_d (h_d / 3) x (elements at
    node i) (the rule is exact for phi^2)."""
    lo, hi = np.array([0.0, 0.25, 0.5][:ndim]), np.array([0.75, 1.25, 1.0][:ndim])
    X, E = fe.box_mesh(ndim, n, lo, hi)
    h = (hi - lo) / np.array(n)
    cnt = _adjacent(E, len(X))
    xi, w = fe.nodal_rule(ndim)
    phi = fe.q1_shape(ndim, xi)
    assert np.array_equal(np.sort(phi.ravel()), np.sort(np.eye(2 ** ndim).ravel()))   # Lagrange: 1 at its node
    M = orc.fe_nodal_mass(E, phi, fe.q1_jxw(X, E, xi, w), len(X))
    assert np.allclose(M, np.prod(h) / 2 ** ndim * cnt, rtol=1e-14, atol=0)
    assert abs(M.sum() - np.prod(hi - lo)) <= 1e-14
    xi, w = fe.gauss_rule(ndim, 2)
    M2 = orc.fe_nodal_mass(E, fe.q1_shape(ndim, xi), fe.q1_jxw(X, E, xi, w), len(X))
    assert np.allclose(M2, np.prod(h / 3.0) * cnt, rtol=1e-13, atol=0)


def test_nodal_spread_conserves_the_integral(orc):
    """sum_n F_n M_n = integral of the Q1 field F (the trapezoid rule is exact for
    multilinear fields on a box mesh): the nodal point source keeps the force."""
    X, E = fe.box_mesh(3, (3, 2, 4), [0.1, 0.2, 0.3], [0.7, 0.5, 0.9])
    xi, w = fe.nodal_rule(3)
    M = orc.fe_nodal_mass(E, fe.q1_shape(3, xi), fe.q1_jxw(X, E, xi, w), len(X))
    A = np.random.default_rng(8).standard_normal((3, 2))
    F = np.array([1.0, -3.0]) + X @ A
    vol = 0.6 * 0.3 * 0.6
    want = vol * (np.array([1.0, -3.0]) + np.array([0.4, 0.35, 0.6]) @ A)
    assert np.allclose((F * M[:, None]).sum(axis=0), want, rtol=1e-13)


# ---- the adaptive quadrature order (getQuadratureKey with get_max_edge_length)

def test_max_edge_length_closed_forms(orc):
    box = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0], [0, 0, 1], [2, 0, 1], [2, 3, 1], [0, 3, 1]], float)
    assert orc.fe_max_edge_length("HEX8", range(8), box) == 3.0        # not the diagonal sqrt(14)
    # HEX20: the general branch takes vertex pairs on edges only; mid-edge nodes moved away are ignored
    X20 = np.concatenate([box, np.full((12, 3), 50.0)])
    assert orc.fe_max_edge_length("HEX20", range(20), X20) == 3.0
    # HEX27: each edge is the sum of its two segments through the mid-edge node
    mids = {8: (0, 1), 9: (1, 2), 10: (2, 3), 11: (0, 3), 12: (0, 4), 13: (1, 5), 14: (2, 6), 15: (3, 7),
            16: (4, 5), 17: (5, 6), 18: (6, 7), 19: (4, 7)}
    X27 = np.zeros((27, 3))
    X27[:8] = box
    for m, (a, b) in mids.items():
        X27[m] = 0.5 * (box[a] + box[b])
    X27[9] += [0.0, 0.0, 2.0]          # the 1-2 edge (length 3) bent by 2 at its middle: 2 * 2.5 = 5
    assert orc.fe_max_edge_length("HEX27", range(27), X27) == 5.0
    # quadrilateral: the longest side, not the longer diagonal
    q = np.array([[0, 0], [4, 0], [4, 3], [0, 3]], float)
    assert orc.fe_max_edge_length("QUAD4", range(4), q) == 4.0
    assert orc.fe_max_edge_length("QUAD9", range(9), np.concatenate([q, np.full((5, 2), 9.0)])) == 4.0
    t = np.array([[0, 0], [3, 0], [0, 4]], float)
    assert orc.fe_max_edge_length("TRI3", range(3), t) == 5.0
    tet = np.array([[0, 0, 0], [3, 0, 0], [0, 4, 0], [0, 0, 12]], float)
    assert orc.fe_max_edge_length("TET4", range(4), tet) == np.sqrt(160.0)   # the 2-3 edge: sqrt(16 + 144)
    t10 = np.zeros((10, 3))
    t10[:4] = tet
    for m, (a, b) in {4: (0, 1), 5: (1, 2), 6: (0, 2), 7: (0, 3), 8: (1, 3), 9: (2, 3)}.items():
        t10[m] = 0.5 * (tet[a] + tet[b])
    t10[4] += [0.0, 2.0, 0.0]         # 0-1 edge (3) bent: 2 * 2.5 = 5 < sqrt(160)
    assert orc.fe_max_edge_length("TET10", range(10), t10) == np.sqrt(40.0) + np.sqrt(40.0)
    t10[8] += [0.0, 20.0, 0.0]        # 1-3 edge: two segments of sqrt(1.5^2 + 20^2 + 6^2) each
    assert orc.fe_max_edge_length("TET10", range(10), t10) == 2 * np.sqrt(1.5 ** 2 + 20 ** 2 + 6 ** 2)
    assert orc.fe_max_edge_length("EDGE2", range(2), np.array([[0, 0, 0], [2, 3, 6]], float)) == 7.0


def test_adaptive_orders_closed_forms(orc):
    """npts = ceil(density * hmax / dx_min), at least 2 (first-order elements) or
    3 (second-order); QGAUSS order min(2 npts - 1, 43), QGRID npts."""
    hexes = []
    X = []
    for k, L in enumerate([0.5, 1.0, 1.25, 30.0]):   # hmax / dx_min = 2 L with dx_min 0.5
        base = len(X)
        X += [[0, 0, 0], [L, 0, 0], [L, L, 0], [0, L, 0], [0, 0, L], [L, 0, L], [L, L, L], [0, L, L]]
        hexes.append(list(range(base, base + 8)))
    X = np.array(X, float)
    g = orc.fe_adaptive_orders("HEX8", hexes, X, 0, 1.0, 0.5)
    # npts: ceil(1) = 1 -> 2; ceil(2) = 2; ceil(2.5) = 3; 60 -> capped order 43
    assert list(g) == [3, 3, 5, 43]
    assert list(orc.fe_adaptive_orders("HEX8", hexes, X, 1, 1.0, 0.5)) == [2, 2, 3, 60]
    assert list(orc.fe_adaptive_orders("HEX8", hexes, X, 0, 2.0, 0.5)) == [3, 7, 9, 43]
    quad9 = [[0, 1, 2, 3, 0, 0, 0, 0, 0]]
    Xq = np.array([[0, 0], [0.5, 0], [0.5, 0.5], [0, 0.5]], float)
    assert list(orc.fe_adaptive_orders("QUAD9", quad9, Xq, 0, 1.0, 0.5)) == [5]     # second order: npts >= 3


# ---- a rule per element

@pytest.mark.parametrize("ndim,n", [(2, (4, 3)), (3, (2, 3, 2))])
def test_rules_per_element(orc, ndim, n):
    """One rule for every element gives the one-rule point source bit for bit;
    mixed rules (1, 2, 3 Gauss points per direction) integrate a Q1 field
    exactly per element, and the right-hand side is the transpose."""
    X, E = fe.box_mesh(ndim, n, [0.1] * ndim, [0.8] * ndim)
    X = X + np.random.default_rng(4).uniform(-0.01, 0.01, X.shape)
    rules = [fe.gauss_rule(ndim, g) for g in (1, 2, 3)]
    phis = [fe.q1_shape(ndim, xi) for xi, _ in rules]
    rng = np.random.default_rng(5)
    F = rng.standard_normal((len(X), 2))
    er = np.full(len(E), 1)
    J1 = fe.q1_jxw(X, E, *rules[1]).reshape(-1)
    Xa, Fa = orc.fe_quadrature_points_rules(E, phis, er, J1, X, F)
    Xb, Fb = orc.fe_quadrature_points(E, phis[1], J1.reshape(len(E), -1), X, F)
    assert np.array_equal(Xa, Xb) and np.array_equal(Fa, Fb)
    er = rng.integers(0, 3, len(E))
    J = np.concatenate([fe.q1_jxw(X, E[e:e + 1], *rules[r]).reshape(-1) for e, r in enumerate(er)])
    A = rng.standard_normal((ndim, 2))
    Flin = np.array([0.5, 2.0]) + X @ A
    Xq, Fq = orc.fe_quadrature_points_rules(E, phis, er, J, X, Flin)
    _, off, _, qp_off = orc.fe_rules_layout(phis, er)
    for e in range(len(E)):
        a, b = qp_off[e], qp_off[e + 1]
        want = orc.fe_quadrature_points(E[e:e + 1], phis[er[e]], J[a:b].reshape(1, -1), X, Flin)[1]
        assert np.array_equal(Fq[a:b], want)
    assert np.allclose(Fq.sum(axis=0), (J.sum() * np.array([0.5, 2.0])) + sum(
        (J[qp_off[e]:qp_off[e + 1]][:, None] * (Xq[qp_off[e]:qp_off[e + 1]] @ A)).sum(axis=0) for e in range(len(E))),
        rtol=1e-13)
    U = rng.standard_normal((int(qp_off[-1]), 2))
    rhs = orc.fe_integrate_rhs_rules(E, phis, er, J, U, len(X))
    _, FJ = orc.fe_quadrature_points_rules(E, phis, er, J, X, F)
    assert abs(np.sum(F * rhs) - np.sum(FJ * U)) <= 1e-12 * np.abs(F * rhs).sum()
