"""Pins the oracle where no reference fixture does: the 1-D weights against
independent closed forms and the defining identities of each kernel, the
spread against the adjoint identity and a brute-force evaluation with the
scalar delta functions, and the binning against exact rational arithmetic.

A plausible slip (a dropped term in the IB_6 polynomial, a wrong sign of r, a
half-cell shift of the stencil, a transposed index) breaks at least one check:
the moments and sums fix every weight, the brute force fixes the placement.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

RNG = np.random.default_rng(20260101)
K6 = 59.0 / 60.0 - math.sqrt(29.0) / 20.0  # Bao-Kaiser-Peskin 6-point kernel second moment


def _dists(orc, kernel, X):
    lo, w = orc.weights_1d(kernel, X, 0.0, 1.0)
    d = np.array([(lo + k + 0.5) - X for k in range(len(w))])
    return d, w


@pytest.mark.parametrize("kernel", ["PIECEWISE_LINEAR", "IB_4", "IB_6", "BSPLINE_3"])
def test_partition_of_unity_and_first_moment(orc, kernel):
    for X in np.concatenate([RNG.uniform(-5, 5, 400), np.arange(-3, 3, 0.25)]):
        d, w = _dists(orc, kernel, X)
        assert abs(w.sum() - 1.0) < 4e-15
        assert abs((d * w).sum()) < 8e-15
        assert np.all(w >= -1e-15)


def test_ib4_matches_scalar_delta(orc):
    """Inline IB_4 weights (lagrangian_interaction3d.f.m4:1266-1273) against the
    scalar delta (lagrangian_delta.f.m4:174-196), a separately written formula."""
    for X in np.concatenate([RNG.uniform(-4, 4, 500), np.arange(-2, 2, 0.125)]):
        d, w = _dists(orc, "IB_4", X)
        for dk, wk in zip(d, w):
            assert abs(wk - orc.ib_4_delta(dk)) < 2e-16 * 8
        # stencil = the 4 cells whose centres are within 2 of X
        assert np.all(np.abs(d) <= 2.0 + 1e-12)


def test_ib4_identities(orc):
    for X in RNG.uniform(-4, 4, 300):
        d, w = _dists(orc, "IB_4", X)
        assert abs(w[0::2].sum() - 0.5) < 4e-15 and abs(w[1::2].sum() - 0.5) < 4e-15
        assert abs((w * w).sum() - 3.0 / 8.0) < 4e-15


def test_ib6_identities(orc):
    """Bao, Kaiser & Peskin (JCP 2016) 6-point kernel: even/odd sums 1/2, moments
    1,0,K,0 with K = 59/60 - sqrt(29)/20, and a constant sum of squares."""
    sq = []
    for X in np.concatenate([RNG.uniform(-4, 4, 400), np.arange(-2, 2, 0.125)]):
        d, w = _dists(orc, "IB_6", X)
        assert np.all(np.abs(d) <= 3.0 + 1e-12)
        assert abs(w[0::2].sum() - 0.5) < 1e-14 and abs(w[1::2].sum() - 0.5) < 1e-14
        assert abs((d * d * w).sum() - K6) < 2e-14
        assert abs((d ** 3 * w).sum()) < 4e-14
        sq.append((w * w).sum())
    assert np.ptp(sq) < 1e-14


def test_ib6_continuity_at_cell_switch(orc):
    """Stencil switch at half-integers (NINT) must be continuous: the weight of a
    given grid point computed just below and just above the switch agrees."""
    for c in range(-3, 4):
        xa = np.nextafter(c + 0.5, -np.inf)
        xb = c + 0.5
        da, wa = _dists(orc, "IB_6", xa)
        db, wb = _dists(orc, "IB_6", xb)
        fa = {round(x - xa, 6): v for x, v in zip(da + xa, wa)}
        fb = {round(x - xb, 6): v for x, v in zip(db + xb, wb)}
        for key in set(fa) | set(fb):
            assert abs(fa.get(key, 0.0) - fb.get(key, 0.0)) < 1e-12


def test_bspline3_closed_form(orc):
    """Quadratic B-spline: 3/4 - d^2 for |d|<=1/2, (3/2-|d|)^2/2 for |d|<=3/2; variance 1/4."""
    for X in RNG.uniform(-4, 4, 400):
        d, w = _dists(orc, "BSPLINE_3", X)
        for dk, wk in zip(d, w):
            a = abs(dk)
            ref = 0.75 - a * a if a <= 0.5 else (0.5 * (1.5 - a) ** 2 if a <= 1.5 else 0.0)
            # the reference evaluates (|x|+1.5)^2 - 6(|x|+1.5) + 9: cancellation costs a few ulp
            assert abs(wk - ref) < 4e-15
        assert abs((d * d * w).sum() - 0.25) < 4e-15


def test_piecewise_linear_closed_form(orc):
    for X in RNG.uniform(-4, 4, 200):
        d, w = _dists(orc, "PIECEWISE_LINEAR", X)
        for dk, wk in zip(d, w):
            assert abs(wk - max(0.0, 1.0 - abs(dk))) < 4e-16


def test_get_cell_index_exact(orc):
    """Cell index against exact rational floor((X - x_lower)/dx); the FP64
    formula may differ only for X within a few ulp of a cell face."""
    n, xl, xu = 64, 0.0, 1.0
    dx = (xu - xl) / n
    pts = list(RNG.uniform(0, 1, 2000)) + [k * dx for k in range(n)]
    pts += [np.nextafter(k * dx, 2.0) for k in range(n)] + [np.nextafter(k * dx, -1.0) for k in range(1, n)]
    X = np.array(pts).reshape(-1, 1).repeat(3, axis=1)
    idx = orc.get_cell_index(X, [xl] * 3, [xu] * 3, [dx] * 3, [0] * 3, [n - 1] * 3)
    for x, i in zip(X[:, 0], idx[:, 0]):
        exact = math.floor((Fraction(x) - Fraction(xl)) / Fraction(dx))
        if exact != i:
            near = min(abs(Fraction(x) - Fraction(k) * Fraction(dx)) for k in (exact, exact + 1))
            assert near < 4 * 2.0 ** -52
    # the x_upper branch: points closer to x_upper use iupper + floor((X-x_upper)/dx) + 1
    nd = 10
    dx10 = 0.1
    x = 0.95
    i = orc.get_cell_index(np.array([[x, x, x]]), [0.0] * 3, [1.0] * 3, [dx10] * 3, [0] * 3, [nd - 1] * 3)
    assert i[0, 0] == 9


def _periodic_fill(u, gw, n):
    """Fill every entry of a (depth, n2, n1, n0) ghosted array outside the first n
    interior entries per axis from the periodic interior (period n cells; for side
    data the upper face n is the image of face 0)."""
    out = u.copy()
    for ax in range(1, u.ndim):
        idx = (np.arange(out.shape[ax]) - gw) % n + gw
        out = np.take(out, idx, axis=ax)
    return out


def _periodic_setup(ndim, n, kernel, npts, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    X = rng.uniform(0, 1, (npts, ndim))
    X[: npts // 4] = rng.uniform(0, 2 * h, (npts // 4, ndim))      # crowd the periodic corner
    X[npts // 4: npts // 2] = 1.0 - rng.uniform(0, 2 * h, (npts // 2 - npts // 4, ndim))
    return X, h


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kernel", ["IB_4", "IB_6", "BSPLINE_3"])
def test_adjoint_identity_cell(orc, ndim, kernel):
    """h^d <S F, u> = <F, J u> on a periodic patch: spread over the ghost-box entry
    list with images (LIndexSetData model), interp over the interior list."""
    n = 12 if ndim == 3 else 24
    X, h = _periodic_setup(ndim, n, kernel, 60, 1)
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl = [0] * ndim, [n - 1] * ndim, [h] * ndim, [0.0] * ndim
    cell = orc.get_cell_index(X, xl, [1.0] * ndim, dx, il, iu)
    F = RNG.standard_normal(len(X))
    sidx, ssh = orc.build_patch_entries(cell, il, iu, gw, [n] * ndim, il, dx, [4] * ndim)
    f = np.zeros(orc.ghosted_shape(il, iu, [gw] * ndim))
    orc.spread(kernel, dx, xl, il, iu, [gw] * ndim, f, X, F, indices=sidx, Xshift=ssh)
    u = np.zeros_like(f)
    core = (slice(None),) + (slice(gw, gw + n),) * ndim
    u[core] = RNG.uniform(-1, 1, u[core].shape)
    u = _periodic_fill(u, gw, n)
    iidx, ish = orc.build_patch_entries(cell, il, iu, gw, [n] * ndim, il, dx, [4] * ndim, interior_only=True)
    assert len(iidx) == len(X)
    V = orc.interp(kernel, dx, xl, il, iu, [gw] * ndim, u, X, indices=iidx, Xshift=ish)
    lhs = h ** ndim * np.sum(f[core] * u[core])
    rhs = np.sum(F * V[:, 0])
    assert abs(lhs - rhs) <= 1e-13 * (abs(rhs) + np.sum(np.abs(F)))


@pytest.mark.parametrize("kernel", ["IB_4", "IB_6", "BSPLINE_3"])
def test_adjoint_identity_side_3d(orc, kernel):
    n = 10
    X, h = _periodic_setup(3, n, kernel, 50, 2)
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl = [0] * 3, [n - 1] * 3, [h] * 3, [0.0] * 3
    cell = orc.get_cell_index(X, xl, [1.0] * 3, dx, il, iu)
    F = RNG.standard_normal((len(X), 3))
    sidx, ssh = orc.build_patch_entries(cell, il, iu, gw, [n] * 3, il, dx, [4] * 3)
    f = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, f, X, F, indices=sidx, Xshift=ssh)
    lhs = 0.0
    us = []
    for a in range(3):
        u = np.zeros((1,) + f[a].shape)
        core = (slice(None),) + (slice(gw, gw + n),) * 3
        u[core] = RNG.uniform(-1, 1, u[core].shape)
        u = _periodic_fill(u, gw, n)
        us.append(np.ascontiguousarray(u[0]))
        lhs += h ** 3 * np.sum(f[a][(slice(gw, gw + n),) * 3] * u[0][(slice(gw, gw + n),) * 3])
        # periodic side data: face 0 and face n along the normal axis agree
        sl0 = [slice(gw, gw + n)] * 3
        sln = list(sl0)
        sl0[2 - a] = gw
        sln[2 - a] = gw + n
        assert np.allclose(f[a][tuple(sl0)], f[a][tuple(sln)], rtol=1e-13, atol=1e-9)
    iidx, ish = orc.build_patch_entries(cell, il, iu, gw, [n] * 3, il, dx, [4] * 3, interior_only=True)
    V = orc.interp_side(kernel, dx, xl, il, iu, [gw] * 3, us, X, indices=iidx, Xshift=ish)
    rhs = np.sum(F * V)
    assert abs(lhs - rhs) <= 1e-13 * (abs(rhs) + np.sum(np.abs(F)))


@pytest.mark.parametrize("kernel", ["IB_4", "BSPLINE_3"])
def test_spread_bruteforce_periodic(orc, kernel):
    """Spread at every interior face of a periodic 3-D side grid equals the
    O(N*G) sum of scalar delta functions (lagrangian_delta.f.m4) over all points
    and periodic images."""
    n = 8
    X, h = _periodic_setup(3, n, kernel, 30, 3)
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl = [0] * 3, [n - 1] * 3, [h] * 3, [0.0] * 3
    cell = orc.get_cell_index(X, xl, [1.0] * 3, dx, il, iu)
    F = RNG.standard_normal((len(X), 3))
    sidx, ssh = orc.build_patch_entries(cell, il, iu, gw, [n] * 3, il, dx, [4] * 3)
    f = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, f, X, F, indices=sidx, Xshift=ssh)
    scale = np.max(np.abs(f[0]))
    for a in range(3):
        for (i0, i1, i2) in [(0, 0, 0), (n - 1, 3, 5), (2, n - 1, 0), (5, 5, 5), (1, 7, 6), (n, 2, 3)]:
            ii = [i0, i1, i2]
            if ii[a] > n or any(ii[d] > n - 1 for d in range(3) if d != a):
                continue
            xg = [(ii[d] + (0.0 if d == a else 0.5)) * h for d in range(3)]
            ref = orc.bruteforce_spread_at(kernel, xg, dx, [1.0] * 3, X, F[:, a])
            got = f[a][i2 + gw, i1 + gw, i0 + gw]
            assert abs(got - ref) <= 1e-12 * scale, (a, ii, got, ref)


def test_spread_order_independence(orc):
    """Any order of the entry list gives the same spread to round-off (the
    reference's node-set iteration order is not fixed by its sources)."""
    n, kernel = 10, "IB_6"
    X, h = _periodic_setup(3, n, kernel, 80, 4)
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl = [0] * 3, [n - 1] * 3, [h] * 3, [0.0] * 3
    cell = orc.get_cell_index(X, xl, [1.0] * 3, dx, il, iu)
    F = RNG.standard_normal((len(X), 3))
    sidx, ssh = orc.build_patch_entries(cell, il, iu, gw, [n] * 3, il, dx, [4] * 3)
    perm = RNG.permutation(len(sidx))
    fa = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    fb = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, fa, X, F, indices=sidx, Xshift=ssh)
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, fb, X, F, indices=sidx[perm], Xshift=ssh[perm])
    for a in range(3):
        assert np.max(np.abs(fa[a] - fb[a])) <= 1e-14 * np.max(np.abs(fa[a]))


def test_side_axis_is_shifted_cell_kernel(orc):
    """Structural identity (LEInteractor.cpp:3679-3710): the side spread along
    axis a equals a cell spread of component a with x_lower shifted by -dx_a/2 on
    the side box."""
    n, kernel = 8, "IB_4"
    X, h = _periodic_setup(3, n, kernel, 20, 5)
    X = 0.3 + 0.4 * X
    gw = 3
    il, iu, dx, xl = [0] * 3, [n - 1] * 3, [h] * 3, [0.0] * 3
    F = RNG.standard_normal((len(X), 3))
    f = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, f, X, F)
    for a in range(3):
        iu_a = list(iu)
        iu_a[a] += 1
        xl_a = list(xl)
        xl_a[a] -= 0.5 * h
        g = np.zeros((1,) + f[a].shape)
        orc.spread(kernel, dx, xl_a, il, iu_a, [gw] * 3, g, X, F[:, a].copy())
        assert np.array_equal(g[0], f[a])


def test_order_entries_z_matches_definition(orc):
    """The spread order of the oracle (group, rank-in-bin // 64, stencil z start,
    rank-in-bin) against a direct O(n^2) evaluation of that definition, with bins
    crowded past one 64-point chunk and periodic images. A point's rank in its
    bin counts the bin's points that come before it by (the cell's z plane
    inside the bin, point index)."""
    rng = np.random.default_rng(41)
    n, N, h, kernel = 700, 16, 1.0 / 16, "IB_6"
    X = rng.uniform(0.0, 1.0, (n, 3))
    X[:300] = rng.uniform(0.0, 3 * h, (300, 3))   # >64 points in the corner bin
    il, iu = [0] * 3, [N - 1] * 3
    cells = orc.get_cell_index(X, [0.0] * 3, [1.0] * 3, [h] * 3, il, iu)
    idx, sh = orc.build_patch_entries(cells, il, iu, 4, [N] * 3, il, [h] * 3, [8] * 3)
    xl = [0.0, 0.0, -0.5 * h]
    perm = orc.order_entries_z(kernel, cells, idx, sh, X, xl, [h] * 3, il, [8] * 3)
    b = cells // 8
    r = cells - 8 * b
    cib = r[:, 2]
    keys = []
    group, prev = -1, None
    for l, p in enumerate(idx):
        o = tuple(int(round(v / h)) for v in sh[l])
        g = (tuple(b[p]), o)
        if g != prev:
            group += 1
            prev = g
        same = np.all(b == b[p], axis=1)
        before = (cib < cib[p]) | ((cib == cib[p]) & (np.arange(n) < p))
        rank = int(np.sum(same & before))
        z0 = orc.weights_1d(kernel, X[p, 2] + sh[l, 2], xl[2], h)[0]
        keys.append((group, rank // 64, z0, rank, l))
    want = [k[-1] for k in sorted(keys)]
    assert list(perm) == want
