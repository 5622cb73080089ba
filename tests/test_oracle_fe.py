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
