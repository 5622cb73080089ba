"""Pins the oracle's BSPLINE_4, BSPLINE_5 and BSPLINE_6 kernels
(lagrangian_delta.f.m4:268-354; stencils lagrangian_interaction3d.f.m4:2882-2908,
:3230-3248, :3460-3480) against the mathematics of cardinal B-splines and
against the one rule the reference adds: BSPLINE_4 and BSPLINE_6 choose the
side of their even stencil by comparing the unshifted X(d,s) with the centre of
the cell of X(d,s)+Xshift, so for a periodic image the shift's sign decides.

A dropped term or wrong coefficient in a polynomial piece breaks the central
values or the moments; a stencil shifted by a cell breaks the support check; a
choice made with the shifted X breaks the image closed forms.  The reference
evaluates each piece as a polynomial in r = |x| + n/2 (terms up to 6^5/120 for
BSPLINE_6 that cancel to the small weights), so the checks allow 1e-12, far
below any such slip (>= 1e-3)."""
from fractions import Fraction

import numpy as np
import pytest

RNG = np.random.default_rng(20261202)
TOL = 1e-12
ORDER = {"BSPLINE_4": 4, "BSPLINE_5": 5, "BSPLINE_6": 6}


def _cardinal(n, x):
    """Cardinal B-spline of order n centred at 0, from its definition
    M_n(x) = 1/(n-1)! sum_k (-1)^k C(n,k) (x + n/2 - k)_+^(n-1), exactly."""
    x = Fraction(x)
    s = Fraction(0)
    fact = 1
    for i in range(2, n):
        fact *= i
    for k in range(n + 1):
        t = x + Fraction(n, 2) - k
        if t > 0:
            c = 1
            for i in range(k):
                c = c * (n - i) // (i + 1)
            s += (-1) ** k * c * t ** (n - 1)
    return s / fact


@pytest.mark.parametrize("kernel", list(ORDER))
def test_delta_equals_cardinal_bspline(orc, kernel):
    n = ORDER[kernel]
    for x in [0, 0.5, 1, 1.5, 2, 2.5, 3] + list(RNG.uniform(-3.2, 3.2, 200)):
        want = float(_cardinal(n, Fraction(float(x))))
        assert abs(orc.delta(kernel, x) - want) < TOL, (x, orc.delta(kernel, x), want)


def test_central_values(orc):
    want = {"BSPLINE_4": [(0, Fraction(2, 3)), (1, Fraction(1, 6))],
            "BSPLINE_5": [(0, Fraction(115, 192)), (1, Fraction(19, 96)), (2, Fraction(1, 384))],
            "BSPLINE_6": [(0, Fraction(11, 20)), (1, Fraction(13, 60)), (2, Fraction(1, 120))]}
    for k, vals in want.items():
        for x, v in vals:
            assert abs(orc.delta(k, x) - float(v)) < TOL
            assert abs(orc.delta(k, -x) - float(v)) < TOL


@pytest.mark.parametrize("kernel", list(ORDER))
def test_moments_and_support(orc, kernel):
    """Sum 1, first moment 0, second moment n/12 (the variance of a sum of n
    uniform variables); every weight is the delta at its cell, and every cell
    within the support is in the stencil."""
    n = ORDER[kernel]
    for X in np.concatenate([RNG.uniform(-5, 5, 400), np.arange(-3, 3, 0.25)]):
        lo, w = orc.weights_1d(kernel, X, 0.0, 1.0)
        assert len(w) == n
        d = np.array([(lo + k + 0.5) - X for k in range(n)])
        assert abs(w.sum() - 1.0) < TOL
        assert abs((d * w).sum()) < TOL
        assert abs((d * d * w).sum() - n / 12.0) < 10 * TOL  # d^2 up to (n/2)^2
        for k in range(n):
            assert w[k] == orc.delta(kernel, -d[k]) or abs(w[k] - orc.delta(kernel, -d[k])) < 1e-16
        c = int(np.floor(X))
        for j in range(c - n, c + n + 1):
            if abs((j + 0.5) - X) < n / 2.0 - 1e-12:
                assert lo <= j < lo + n, (X, j, lo)


@pytest.mark.parametrize("kernel", ["BSPLINE_4", "BSPLINE_6"])
def test_image_side_chosen_by_unshifted_x(orc, kernel):
    """An image shifted by +L has X(d,s) < centre always (lower side), one shifted
    by -L the upper side.  When that is the other side than X+Xshift alone would
    pick, the stencil drops the far support cell: the weights sum to
    1 - M_n(n/2 - |t|) = 1 - |t|^(n-1)/(n-1)!, t the offset from the centre."""
    n = ORDER[kernel]
    fact = 6 if n == 4 else 120
    L = 16.0
    for _ in range(300):
        c = int(RNG.integers(-4, 4))
        t = float(RNG.uniform(-0.5, 0.5))
        Xs = c + 0.5 + t
        for shift in (L, -L):
            lo, w = orc.weights_1d(kernel, Xs, 0.0, 1.0, Xu=Xs - shift)
            forced_lower = shift > 0
            assert lo == (c - n // 2 if forced_lower else c - n // 2 + 1)
            natural_lower = t < 0
            if forced_lower == natural_lower:
                assert abs(w.sum() - 1.0) < TOL
            else:
                assert abs(w.sum() - (1.0 - abs(t) ** (n - 1) / fact)) < TOL, (t, shift, w.sum())
        # no shift: the natural side
        lo0, w0 = orc.weights_1d(kernel, Xs, 0.0, 1.0)
        assert lo0 == (c - n // 2 if t < 0 else c - n // 2 + 1)


@pytest.mark.parametrize("kernel", list(ORDER))
def test_spread_bruteforce(orc, kernel):
    """Away from the periodic boundary (no images) the 3-D side spread equals the
    O(N*G) sum of the scalar deltas; BSPLINE_5 (centred stencil) also with images."""
    n = 8
    h = 1.0 / n
    rng = np.random.default_rng(7)
    X = rng.uniform(0.35, 0.65, (20, 3)) if kernel != "BSPLINE_5" else rng.uniform(0, 1, (20, 3))
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl = [0] * 3, [n - 1] * 3, [h] * 3, [0.0] * 3
    cell = orc.get_cell_index(X, xl, [1.0] * 3, dx, il, iu)
    F = rng.standard_normal((len(X), 3))
    sidx, ssh = orc.build_patch_entries(cell, il, iu, gw, [n] * 3, il, dx, [4] * 3)
    f = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, f, X, F, indices=sidx, Xshift=ssh)
    scale = np.max(np.abs(f[0]))
    for a in range(3):
        for (i0, i1, i2) in [(4, 4, 4), (3, 5, 4), (2, 4, 6), (5, 3, 3), (0, 0, 0), (7, 1, 6)]:
            ii = [i0, i1, i2]
            xg = [(ii[d] + (0.0 if d == a else 0.5)) * h for d in range(3)]
            ref = orc.bruteforce_spread_at(kernel, xg, dx, [1.0] * 3, X, F[:, a])
            got = f[a][i2 + gw, i1 + gw, i0 + gw]
            assert abs(got - ref) <= 1e-12 * scale, (a, ii, got, ref)
