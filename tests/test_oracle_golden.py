"""Pins the CPU oracle against the reference's own golden outputs.

Fixtures (tests/golden/*.json) hold the data printed by the reference tests:
  tests/interpolate/interpolate_01.cpp (points: mt19937(42), uniform over the
  fine patch, :160-177; field filled at cell centres, :121-141) and
  tests/IBTK/ghost_accumulation_01.cpp (Q then X drawn from mt19937(42), :131-135;
  spread with the raw-array overload, :137-156).
The point sets are regenerated bit-for-bit (tests/golden/mt19937.py), so the
comparison is against exact inputs.  Tolerances are print precision (16
significant digits): 2e-15 relative, except for the BSPLINE_4/5/6 kernels,
whose polynomial form bounds the agreement between two compilations (COND_TOL).
"""
import json
import math
import os

import numpy as np
import pytest

from mt19937 import MT19937, uniform_real

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KERNELS = [("IB_4", "ib_4"), ("IB_6", "ib_6"), ("BSPLINE_3", "bspline_3"), ("BSPLINE_4", "bspline_4"),
           ("BSPLINE_5", "bspline_5"), ("BSPLINE_6", "bspline_6")]
PRINT_TOL = 2e-15
# BSPLINE_4/5/6 evaluate each piece as a polynomial in r = |x| + n/2 whose terms
# (up to 64/6, 625/24 and 7776/120) cancel to weights of order 1e-2, so one
# rounding difference in a term (the reference's compiler may contract or
# reorder; the oracle follows the source order with no contraction) moves a
# weight by ~1e-14 and V by up to ~1e-12 relative.  Any slip in a coefficient
# or stencil is >= 1e-4.
COND_TOL = {"BSPLINE_4": 5e-14, "BSPLINE_5": 5e-13, "BSPLINE_6": 3e-12}


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_mt19937_known_answer():
    # ISO C++ [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995.
    g = MT19937()
    for _ in range(9999):
        g()
    assert g() == 4123659995


def _points(n, ndim, lo, hi):
    g = MT19937(42)
    return np.array([uniform_real(g, lo, hi) for _ in range(n * ndim)]).reshape(n, ndim)


@pytest.mark.parametrize("kernel,tag", KERNELS)
def test_interpolate_01_3d_golden(orc, kernel, tag):
    """interpolate_01_3d.<kernel>: coarse 8^3, fine patch (8..15)^3, dx=1/32, trilinear field."""
    rows = np.array(_load("interpolate_01_3d.%s.json" % tag)["rows"])
    X = _points(100, 3, 0.25, 0.5)
    assert np.max(np.abs(X - rows[:, :3]) / np.abs(rows[:, :3])) < 1e-15
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl, xu = [8] * 3, [15] * 3, [1 / 32] * 3, [0.25] * 3, [0.5] * 3
    k = np.arange(-gw, 8 + gw)
    xc = xl[0] + (k + 0.5) * dx[0]
    Z, Y, Xg = np.meshgrid(xc, xc, xc, indexing="ij")
    f = 1 + 2 * Xg + 3 * Y - Z + 4 * Xg * Y + 2 * Xg * Z + 3 * Xg * Y * Z
    u = np.ascontiguousarray(np.broadcast_to(f, (3,) + f.shape))
    idx = orc.build_local_indices(il, iu, xl, xu, dx, il, iu, X)
    assert len(idx) == 100
    V = orc.interp(kernel, dx, xl, il, iu, [gw] * 3, u, X, indices=idx, depth=3)
    for c in range(3):
        err = np.abs(V[:, c] - rows[:, 3 + c]) / np.abs(rows[:, 3 + c])
        assert err.max() < COND_TOL.get(kernel, PRINT_TOL), (kernel, c, err.max())


@pytest.mark.parametrize("kernel,tag", KERNELS)
def test_interpolate_01_2d_golden_interior(orc, kernel, tag):
    """interpolate_01_2d.<kernel>: coarse 16^2, fine patch (16..31)^2, dx=1/64,
    u = sin(2 pi (x-.1234)) sin(2 pi (y-.1234)).  Ghost values come from SAMRAI
    refinement, which is not restated, so only points whose whole stencil lies in
    the patch interior are compared; ghost cells are NaN so any use would show."""
    rows = np.array(_load("interpolate_01_2d.%s.json" % tag)["rows"])
    X = _points(100, 2, 0.25, 0.5)
    assert np.max(np.abs(X - rows[:, :2]) / np.abs(rows[:, :2])) < 1e-15
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl = [16] * 2, [31] * 2, [1 / 64] * 2, [0.25] * 2
    k = np.arange(-gw, 16 + gw)
    xc = xl[0] + (k + 0.5) * dx[0]
    Y, Xg = np.meshgrid(xc, xc, indexing="ij")
    f = np.sin(2 * math.pi * (Xg - 0.1234)) * np.sin(2 * math.pi * (Y - 0.1234))
    inner = np.zeros(f.shape, dtype=bool)
    inner[gw:gw + 16, gw:gw + 16] = True
    f = np.where(inner, f, np.nan)
    u = np.ascontiguousarray(np.broadcast_to(f, (2,) + f.shape))
    V = orc.interp(kernel, dx, xl, il, iu, [gw] * 2, u, X, depth=2)
    ok = ~np.isnan(V[:, 0])
    assert ok.sum() >= 40
    for c in range(2):
        err = np.abs(V[ok, c] - rows[ok, 3 + c]) / np.abs(rows[ok, 3 + c])
        assert err.max() < COND_TOL.get(kernel, PRINT_TOL), (kernel, c, err.max())


def test_ghost_accumulation_3d_side_spread_golden(orc):
    """ghost_accumulation_01_3d.side.spread.a: one non-periodic 8^3 patch, side
    data, ghost width 2, PIECEWISE_LINEAR.  Pins x_lower - dx/2 per axis,
    toSideBox and the per-axis scalar split of the side-data spread."""
    fx = _load("ghost_accumulation_01_3d.side.spread.a.json")
    g = MT19937(42)
    Q = np.array([uniform_real(g, 0.0, 1.0) for _ in range(300)]).reshape(100, 3)
    X = np.array([uniform_real(g, 0.0, 1.0) for _ in range(300)]).reshape(100, 3)
    il, iu, gcw, dx, xl, xu = [0] * 3, [7] * 3, [2] * 3, [1 / 8] * 3, [0.0] * 3, [1.0] * 3
    qa = [np.zeros(orc.side_axis_shape(il, iu, gcw, a)) for a in range(3)]
    idx = orc.build_local_indices(il, iu, xl, xu, dx, il, iu, X)
    orc.spread_side("PIECEWISE_LINEAR", dx, xl, il, iu, gcw, qa, X, Q, indices=idx)
    n = 0
    for a in range(3):
        for i0, i1, i2, v in fx["axes"][str(a)]:
            got = qa[a][i2 + 2, i1 + 2, i0 + 2]
            if v == 0.0:
                assert got == 0.0
            else:
                assert abs(got - v) <= PRINT_TOL * abs(v), (a, i0, i1, i2, got, v)
            n += 1
    assert n == 3 * 9 * 8 * 8


@pytest.mark.parametrize("case", ["2d.cell.spread.a", "2d.cell.spread.b", "2d.side.spread.a", "2d.side.spread.b",
                                  "3d.cell.spread.a", "3d.cell.spread.b", "3d.side.spread.a", "3d.side.spread.b"])
def test_ghost_accumulation_all_cases_golden(orc, case):
    """tests/IBTK/ghost_accumulation_01 spread cases (PIECEWISE_LINEAR, one or
    several patches, periodic or not, cell and side data): the reference spreads
    each patch's points over the patch, then SAMRAIGhostDataAccumulator sums every
    degree of freedom over all patches' (ghosted) copies.  The oracle instead
    spreads, for each patch, every point of the level within the patch's index
    ghost box (with periodic images) -- the LDataManager::spread model the GPU
    level plans use.  Both must give the same patch values (to summation order)."""
    import ghost_cases as gc
    c = gc.load(case)
    keep = gc.level_points(orc, c)
    X, Q = c["X"][keep], c["Q"][keep]
    ndim, h = c["ndim"], c["h"]
    cells = orc.get_cell_index(X, [0.0] * ndim, [1.0] * ndim, [h] * ndim, [0] * ndim, [c["nf"] - 1] * ndim)
    n = 0
    for pi, (lo, hi, xl, xu) in enumerate(c["boxes"]):
        if c["side"]:
            arrays = [np.zeros(orc.side_axis_shape(lo, hi, c["gcw"], a)) for a in range(ndim)]
        else:
            arrays = np.zeros(orc.ghosted_shape(lo, hi, c["gcw"], depth=Q.shape[1]))
        orc.spread_patch(c["kernel"], c["side"], [h] * ndim, xl, lo, hi, c["gcw"], arrays, X, Q, cells,
                         c["gcw"][0], c["pshift"], [0] * ndim, depth=Q.shape[1])
        n += gc.compare(c, pi, arrays if c["side"] else [arrays])
    assert n > 0
