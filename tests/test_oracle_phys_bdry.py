"""The oracle's restatement of the LINEAR Robin boundary operators
(cartphysbdryop{2,3}d.f.m4 via CartCellRobinPhysBdryOp /
CartSideRobinPhysBdryOp) pinned by what they are: the accumulation
(adjoint_op = 1) is the transpose of the homogeneous ghost fill on the
interior values, <fold(f), u> = <f, fill(u)>, for every codimension; and the
fill reproduces the closed forms of Dirichlet / Neumann linear extrapolation."""
import numpy as np
import pytest

import ibamr_amd as ia


def _patch_level(ndim, n=6, periodic=()):
    h = 1.0 / n
    p = ia.Patch([0] * ndim, [n - 1] * ndim, [0.0] * ndim, [1.0] * ndim)
    return ia.Level([h] * ndim, [0] * ndim, [n - 1] * ndim, [0.0] * ndim, [1.0] * ndim,
                    [n if d in periodic else 0 for d in range(ndim)], [p])


def _shape(ndim, n, gcw, side, axis, depth):
    s = [n + 2 * gcw + (1 if side and d == axis else 0) for d in range(ndim)]
    s = tuple(reversed(s))
    return s if side else (depth,) + s


def _interior_mask(ndim, n, gcw, side, axis, depth):
    m = np.zeros(_shape(ndim, n, gcw, side, axis, depth), dtype=bool)
    sl = tuple(slice(gcw, gcw + n + (1 if side and (ndim - 1 - k) == axis else 0)) for k in range(ndim))
    if side:
        m[sl] = True
    else:
        m[(slice(None),) + sl] = True
    return m


def _coefs(rng, nboxes, narr, kinds):
    abg = np.zeros((nboxes, narr, 3))
    for k in range(nboxes):
        for a in range(narr):
            kind = kinds[rng.integers(0, len(kinds))]
            abg[k, a] = {"dirichlet": (1.0, 0.0, 0.0), "neumann": (0.0, 1.0, 0.0),
                         "robin": (rng.uniform(0.5, 2.0), rng.uniform(0.1, 1.0), 0.0)}[kind]
    return abg


@pytest.mark.parametrize("ndim,side,kinds", [
    (2, False, ("dirichlet", "neumann", "robin")), (3, False, ("dirichlet", "neumann", "robin")),
    (2, True, ("neumann", "robin")), (3, True, ("neumann", "robin"))])
def test_fold_is_transpose_of_fill(orc, ndim, side, kinds):
    n, gcw, depth = 6, 3, (1 if side else 2)
    level = _patch_level(ndim, n)
    boxes = ia.physical_boundary_boxes(level, 0, gcw)
    assert sum(1 for b in boxes if b[0] == 1) == 2 * ndim
    assert sum(1 for b in boxes if b[0] == ndim) == 2 ** ndim
    rng = np.random.default_rng(ndim * 10 + side)
    narr = ndim if side else depth
    abg = _coefs(rng, len(boxes), narr, kinds)
    axes = range(ndim) if side else [0]
    masks = [_interior_mask(ndim, n, gcw, side, a, depth) for a in axes]
    u = [np.where(m, rng.standard_normal(m.shape), 0.0) for m in masks]
    f = [rng.standard_normal(m.shape) for m in masks]
    Mu = orc.phys_bdry([x.copy() for x in u], [0] * ndim, [n - 1] * ndim, gcw, level.dx, boxes, abg, side, False,
                       depth=depth)
    Af = orc.phys_bdry([x.copy() for x in f], [0] * ndim, [n - 1] * ndim, gcw, level.dx, boxes, abg, side, True,
                       depth=depth)
    lhs = sum(float((a[m] * b[m]).sum()) for a, b, m in zip(Af, u, masks))
    rhs = sum(float((a * b).sum()) for a, b in zip(f, Mu))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (lhs, rhs)
    # the fill leaves the interior alone except a Dirichlet normal face (none here)
    for x, y, m in zip(u, Mu, masks):
        assert np.array_equal(x[m], y[m])


def test_cell_fill_closed_forms(orc):
    """Dirichlet: ghost = -u_mirror + 2 g; Neumann: ghost = u_mirror + (1 + 2i) h g."""
    n, gcw, ndim = 6, 3, 2
    level = _patch_level(ndim, n, periodic=(1,))
    boxes = ia.physical_boundary_boxes(level, 0, gcw)
    assert [b[:2] for b in boxes] == [(1, 0), (1, 1)]
    abg = np.array([[[1.0, 0.0, 0.7]], [[0.0, 1.0, -0.4]]])
    rng = np.random.default_rng(3)
    u = rng.standard_normal(_shape(ndim, n, gcw, False, 0, 1))
    out = orc.phys_bdry([u.copy()], [0, 0], [n - 1, n - 1], gcw, level.dx, boxes, abg, False, False)[0]
    h = 1.0 / n
    for i in range(gcw):
        lo_g, lo_i = gcw - 1 - i, gcw + i
        hi_g, hi_i = gcw + n + i, gcw + n - 1 - i
        assert np.allclose(out[0, gcw:gcw + n, lo_g], -u[0, gcw:gcw + n, lo_i] + 2 * 0.7, atol=1e-14)
        assert np.allclose(out[0, gcw:gcw + n, hi_g], u[0, gcw:gcw + n, hi_i] + (1 + 2 * i) * h * -0.4, atol=1e-14)


def test_side_dirichlet_normal_fold_closed_form(orc):
    """Normal component, Dirichlet: the boundary face takes g/a plus twice each
    ghost value, the mirrored interior faces lose the ghost values."""
    n, gcw, ndim = 5, 2, 2
    level = _patch_level(ndim, n, periodic=(1,))
    boxes = ia.physical_boundary_boxes(level, 0, gcw)
    abg = np.zeros((len(boxes), 2, 3))
    abg[:, 0] = (2.0, 0.0, 0.6)     # normal (x) component: Dirichlet g/a = 0.3
    abg[:, 1] = (0.0, 1.0, 0.0)     # tangential: Neumann
    rng = np.random.default_rng(4)
    f = [rng.standard_normal(_shape(ndim, n, gcw, True, a, 1)) for a in range(2)]
    out = orc.phys_bdry([x.copy() for x in f], [0, 0], [n - 1, n - 1], gcw, level.dx, boxes, abg, True, True)
    u0, g0 = out[0], f[0]
    rows = slice(gcw, gcw + n)
    ib = gcw  # lower boundary face x = 0
    want_face = 0.3 + 2 * g0[rows, ib - 1] + 2 * g0[rows, ib - 2]
    assert np.allclose(u0[rows, ib], want_face, atol=1e-14)
    assert np.allclose(u0[rows, ib + 1], g0[rows, ib + 1] - g0[rows, ib - 1], atol=1e-14)
    assert np.allclose(u0[rows, ib + 2], g0[rows, ib + 2] - g0[rows, ib - 2], atol=1e-14)


def _varying(rng, boxes, narr, ndim, side, kinds, g_scale=0.0):
    """Coefficient arrays varying over each codim-1 box (every column its own
    kind among `kinds` and its own values), in the orc_phys_bdry_coefs layout."""
    out = []
    for b in boxes:
        if b[0] != 1:
            continue
        for j in range(narr):
            m = orc_ncoef(ndim, b, j, side)
            a, bb, g = np.zeros(m), np.zeros(m), g_scale * rng.standard_normal(m)
            for i in range(m):
                kind = kinds[rng.integers(0, len(kinds))]
                a[i], bb[i] = {"dirichlet": (1.0, 0.0), "neumann": (0.0, 1.0),
                               "robin": (rng.uniform(0.5, 2.0), rng.uniform(0.1, 1.0))}[kind]
            out += [a, bb, g]
    return np.concatenate(out) if out else np.zeros(0)


def orc_ncoef(ndim, box, j, side):
    import oracle
    return oracle.phys_bdry_ncoef(ndim, box, j, side)


@pytest.mark.parametrize("ndim,side", [(2, False), (3, False), (2, True), (3, True)])
def test_varying_coefficients_constant_equal_uniform(orc, ndim, side):
    """Coefficient arrays holding one value per box and array give the uniform
    form's bits (fold and fill)."""
    n, gcw, depth = 6, 3, (1 if side else 2)
    level = _patch_level(ndim, n)
    boxes = ia.physical_boundary_boxes(level, 0, gcw)
    rng = np.random.default_rng(ndim * 7 + side)
    narr = ndim if side else depth
    abg = _coefs(rng, len(boxes), narr, ("neumann", "robin") if side else ("dirichlet", "neumann", "robin"))
    abg[:, :, 2] = rng.standard_normal((len(boxes), narr))
    coefs = []
    for k, b in enumerate(boxes):
        if b[0] == 1:
            for j in range(narr):
                m = orc_ncoef(ndim, b, j, side)
                coefs += [np.full(m, abg[k, j, 0]), np.full(m, abg[k, j, 1]), np.full(m, abg[k, j, 2])]
    coefs = np.concatenate(coefs)
    axes = range(ndim) if side else [0]
    for adjoint in (False, True):
        u = [rng.standard_normal(_shape(ndim, n, gcw, side, a, depth)) for a in axes]
        want = orc.phys_bdry([x.copy() for x in u], [0] * ndim, [n - 1] * ndim, gcw, level.dx, boxes, abg, side,
                             adjoint, depth=depth)
        got = orc.phys_bdry_coefs([x.copy() for x in u], [0] * ndim, [n - 1] * ndim, gcw, level.dx, boxes, coefs,
                                  side, adjoint, depth=depth)
        for x, y in zip(got, want):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("ndim,side,kinds", [
    (2, False, ("dirichlet", "neumann", "robin")), (3, False, ("dirichlet", "neumann", "robin")),
    (2, True, ("neumann", "robin")), (3, True, ("neumann", "robin"))])
def test_varying_coefficients_fold_is_transpose_of_fill(orc, ndim, side, kinds):
    n, gcw, depth = 6, 3, (1 if side else 2)
    level = _patch_level(ndim, n)
    boxes = ia.physical_boundary_boxes(level, 0, gcw)
    rng = np.random.default_rng(ndim * 13 + side)
    narr = ndim if side else depth
    coefs = _varying(rng, boxes, narr, ndim, side, kinds)
    axes = range(ndim) if side else [0]
    masks = [_interior_mask(ndim, n, gcw, side, a, depth) for a in axes]
    u = [np.where(m, rng.standard_normal(m.shape), 0.0) for m in masks]
    f = [rng.standard_normal(m.shape) for m in masks]
    args = ([0] * ndim, [n - 1] * ndim, gcw, level.dx, boxes, coefs, side)
    Mu = orc.phys_bdry_coefs([x.copy() for x in u], *args, False, depth=depth)
    Af = orc.phys_bdry_coefs([x.copy() for x in f], *args, True, depth=depth)
    lhs = sum(float((a[m] * b[m]).sum()) for a, b, m in zip(Af, u, masks))
    rhs = sum(float((a * b).sum()) for a, b in zip(f, Mu))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (lhs, rhs)


def test_varying_coefficients_closed_forms_per_cell(orc):
    """3-D cell data, x faces: at the lower face Dirichlet with g varying over
    the face (ghost = -u_mirror + 2 g(y, z)), at the upper face Neumann with g
    varying (ghost = u_mirror + (1 + 2i) h g(y, z)); the coefficient arrays run
    y fastest (the Fortran acoef(j, k) order), which a transposed reading
    fails."""
    n, gcw, ndim = 5, 2, 3
    level = _patch_level(ndim, n, periodic=(1, 2))
    boxes = ia.physical_boundary_boxes(level, 0, gcw)
    assert [b[:2] for b in boxes] == [(1, 0), (1, 1)]
    rng = np.random.default_rng(8)
    gl = rng.standard_normal((n, n))   # [z][y]
    gu = rng.standard_normal((n, n))
    one, zero = np.ones(n * n), np.zeros(n * n)
    coefs = np.concatenate([one, zero, gl.reshape(-1), zero, one, gu.reshape(-1)])
    u = rng.standard_normal(_shape(ndim, n, gcw, False, 0, 1))
    out = orc.phys_bdry_coefs([u.copy()], [0] * 3, [n - 1] * 3, gcw, level.dx, boxes, coefs, False, False)[0]
    h = 1.0 / n
    c = slice(gcw, gcw + n)
    for i in range(gcw):
        lo_g, lo_i = gcw - 1 - i, gcw + i
        hi_g, hi_i = gcw + n + i, gcw + n - 1 - i
        assert np.allclose(out[0, c, c, lo_g], -u[0, c, c, lo_i] + 2 * gl, atol=1e-14)
        assert np.allclose(out[0, c, c, hi_g], u[0, c, c, hi_i] + (1 + 2 * i) * h * gu, atol=1e-14)
