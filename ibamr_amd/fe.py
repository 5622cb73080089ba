"""Host-side finite-element tables for the FE point source
(ibamr_fe_quadrature_points): what libMesh supplies to FEDataManager::spread in
the reference (shape function values at the quadrature points of a rule, and
JxW per element), for the first-order Lagrange elements QUAD4 and HEX8 with
tensor-product Gauss-Legendre rules and the nodal (trapezoid) rule, plus
structured box meshes.

Node numbering follows libMesh's QUAD4 / HEX8 convention: reference
coordinates in [-1, 1]^d, nodes counter-clockwise in the bottom face first
(HEX8: nodes 4-7 are nodes 0-3 lifted to xi_2 = +1)."""
import itertools

import numpy as np

_CORNERS = {2: np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=np.float64),
            3: np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                         [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)}


def gauss_rule(ndim, npts):
    """Tensor Gauss-Legendre rule with npts points per direction on [-1, 1]^ndim:
    (points [nqp][ndim], weights [nqp]), the first coordinate fastest."""
    x, w = np.polynomial.legendre.leggauss(npts)
    pts, wts = [], []
    for idx in itertools.product(range(npts), repeat=ndim):
        i = idx[::-1]  # first coordinate fastest
        pts.append([x[k] for k in i])
        wts.append(np.prod([w[k] for k in i]))
    return np.array(pts), np.array(wts)


def nodal_rule(ndim):
    """The nodal (trapezoid) rule of first-order Lagrange elements, libMesh's
    build_nodal_qrule for QUAD4 / HEX8: the 2^ndim corners of [-1, 1]^ndim,
    first coordinate fastest, weight 1 each."""
    pts = np.array([[(-1.0, 1.0)[(i >> d) & 1] for d in range(ndim)] for i in range(2 ** ndim)])
    return pts, np.ones(2 ** ndim)


def gauss_rule_for_order(ndim, order):
    """The tensor Gauss rule libMesh's QGAUSS builds for a given order: order // 2 + 1
    points per direction (exact for polynomials of that degree per direction)."""
    return gauss_rule(ndim, order // 2 + 1)


def q1_shape(ndim, xi):
    """Q1 shape functions phi[k][q] at reference points xi [nqp][ndim]."""
    c = _CORNERS[ndim]
    phi = np.ones((len(c), len(xi)))
    for k in range(len(c)):
        for d in range(ndim):
            phi[k] *= 0.5 * (1.0 + c[k, d] * xi[:, d])
    return phi


def q1_shape_grad(ndim, xi):
    """d phi[k][q] / d xi_d as [ndim][nbasis][nqp]."""
    c = _CORNERS[ndim]
    g = np.ones((ndim, len(c), len(xi)))
    for j in range(ndim):
        for k in range(len(c)):
            for d in range(ndim):
                g[j, k] *= 0.5 * c[k, d] if d == j else 0.5 * (1.0 + c[k, d] * xi[:, d])
    return g


def q1_jxw(X_nodes, elem_nodes, xi, w):
    """JxW[e][q] = det(dX/dxi) * w_q for Q1 elements (as libMesh's FE map)."""
    ndim = X_nodes.shape[1]
    g = q1_shape_grad(ndim, xi)                       # [ndim][nb][nqp]
    Xe = X_nodes[elem_nodes]                          # [ne][nb][ndim]
    J = np.einsum("jkq,ekd->eqdj", g, Xe)             # dX_d / dxi_j
    return np.linalg.det(J) * w[None, :]


def box_mesh(ndim, n, lower, upper):
    """Structured Q1 mesh of the box: (X_nodes [nnodes][ndim], elem_nodes [nelem][2^ndim] int32)."""
    lower, upper = np.asarray(lower, float), np.asarray(upper, float)
    axes = [np.linspace(lower[d], upper[d], n[d] + 1) for d in range(ndim)]
    grid = np.meshgrid(*axes, indexing="ij")
    X = np.stack([g.ravel(order="F") for g in grid], axis=1)   # first index fastest

    def nid(ijk):
        s, m = 0, 1
        for d in range(ndim):
            s += ijk[d] * m
            m *= n[d] + 1
        return s
    elems = []
    for e in itertools.product(*[range(n[d]) for d in reversed(range(ndim))]):
        e = e[::-1]
        elems.append([nid([e[d] + (1 if _CORNERS[ndim][k, d] > 0 else 0) for d in range(ndim)])
                      for k in range(2 ** ndim)])
    return X, np.array(elems, dtype=np.int32)
