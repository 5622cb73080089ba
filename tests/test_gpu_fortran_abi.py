"""GPU: the link-level drop-in for IBTK's Fortran kernels (S0,
include/ibamr_amd_fortran.h) against the CPU oracle and the reference's golden
vectors.  The S0 path sums spread contributions in list order, so both
interp and spread must be bitwise equal to the oracle's reference loops for
any entry order, including repeated indices and stencils clipped at the
ghost box."""
import ctypes
import json
import os

import numpy as np
import pytest

import ibamr_amd as ia
from mt19937 import MT19937, uniform_real

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _case(ndim, n, gw, npts, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    il, iu = [2] * ndim, [2 + n - 1] * ndim
    xl = [2 * h] * ndim
    # points in and around the patch (some stencils clipped by the ghost box)
    X = rng.uniform(xl[0] - (gw + 1.5) * h, xl[0] + (n + gw + 1.5) * h, (npts, ndim))
    idx = rng.permutation(npts).astype(np.int32)
    idx = np.concatenate([idx, idx[: npts // 5]])              # repeated indices
    sh = rng.integers(-1, 2, (len(idx), ndim)) * (n * h)        # image shifts
    return rng, h, il, iu, xl, X, idx, sh


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kernel", ["IB_4", "IB_6", "BSPLINE_3", "PIECEWISE_LINEAR", "BSPLINE_4", "BSPLINE_5",
                                    "BSPLINE_6"])
@pytest.mark.parametrize("depth", [1, 2])
def test_s0_bitwise(orc, ndim, kernel, depth):
    print("s0", ndim, kernel, depth, flush=True)
    n = 10 if ndim == 3 else 24
    gw = ia.get_minimum_ghost_width(kernel)
    rng, h, il, iu, xl, X, idx, sh = _case(ndim, n, gw, 150, 17 + depth)
    shp = orc.ghosted_shape(il, iu, [gw] * ndim, depth)
    u = rng.uniform(-1, 1, shp)
    V0 = rng.uniform(-1, 1, (X.shape[0], depth))
    # interp
    Vg = V0.copy()
    ia.s0_interp(kernel, [h] * ndim, xl, il, iu, [gw] * ndim, u, X, Vg, idx, sh, depth=depth)
    Vo = orc.interp(kernel, [h] * ndim, xl, il, iu, [gw] * ndim, u, X, indices=idx, Xshift=sh, depth=depth)
    # oracle leaves untouched rows NaN; the drop-in leaves them as they were
    touched = ~np.isnan(Vo[:, 0])
    assert np.array_equal(Vg[touched], Vo[touched])
    assert np.array_equal(Vg[~touched], V0[~touched])
    # spread
    F = rng.standard_normal((X.shape[0], depth))
    ug = u.copy()
    ia.s0_spread(kernel, [h] * ndim, xl, il, iu, [gw] * ndim, ug, X, F, idx, sh, depth=depth)
    uo = u.copy()
    orc.spread(kernel, [h] * ndim, xl, il, iu, [gw] * ndim, uo, X, F, indices=idx, Xshift=sh, depth=depth)
    assert np.array_equal(ug, uo)


@pytest.mark.parametrize("kernel,tag", [("IB_4", "ib_4"), ("IB_6", "ib_6"), ("BSPLINE_3", "bspline_3"),
                                        ("BSPLINE_4", "bspline_4"), ("BSPLINE_5", "bspline_5"),
                                        ("BSPLINE_6", "bspline_6")])
def test_s0_interpolate_01_3d_golden(kernel, tag):
    """The reference test interpolate_01_3d.<kernel> through the Fortran-ABI
    drop-in on the GPU: 16-digit agreement with the printed values."""
    with open(os.path.join(GOLD, "interpolate_01_3d.%s.json" % tag)) as f:
        rows = np.array(json.load(f)["rows"])
    g = MT19937(42)
    X = np.array([uniform_real(g, 0.25, 0.5) for _ in range(300)]).reshape(100, 3)
    gw = ia.get_minimum_ghost_width(kernel)
    k = np.arange(-gw, 8 + gw)
    xc = 0.25 + (k + 0.5) / 32
    Z, Y, Xg = np.meshgrid(xc, xc, xc, indexing="ij")
    fld = 1 + 2 * Xg + 3 * Y - Z + 4 * Xg * Y + 2 * Xg * Z + 3 * Xg * Y * Z
    u = np.ascontiguousarray(np.broadcast_to(fld, (3,) + fld.shape))
    V = np.zeros((100, 3))
    lib = ia.lib()
    idx = np.arange(100, dtype=np.int32)
    ia.s0_interp(kernel, [1 / 32] * 3, [0.25] * 3, [8] * 3, [15] * 3, [gw] * 3, u, X, V, idx, np.zeros((100, 3)),
                 depth=3)
    for c in range(3):
        err = np.abs(V[:, c] - rows[:, 3 + c]) / np.abs(rows[:, 3 + c])
        # print precision; BSPLINE_4/5/6: the conditioning of their polynomial form (test_oracle_golden)
        assert err.max() < {"BSPLINE_4": 5e-14, "BSPLINE_5": 5e-13, "BSPLINE_6": 3e-12}.get(kernel, 2e-15)


def test_s0_fortran_symbol_call():
    """Call lagrangian_ib_4_interp3d_ exactly as LEInteractor.cpp:4673-4701 does
    (every scalar by reference)."""
    lib = ia.lib()
    f = lib.lagrangian_ib_4_interp3d_
    n, gw, depth = 6, 3, 1
    h = 1.0 / n
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, (depth, n + 2 * gw, n + 2 * gw, n + 2 * gw))
    X = rng.uniform(0.2, 0.8, (7, 3))
    V = np.zeros((7, depth))
    idx = np.arange(7, dtype=np.int32)
    sh = np.zeros((7, 3))
    dxa = np.array([h] * 3)
    xla = np.zeros(3)
    xua = np.ones(3)
    ci = ctypes.c_int
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    R = lambda v: ctypes.byref(ci(v))
    f.restype = None
    f(P(dxa, ctypes.c_double), P(xla, ctypes.c_double), P(xua, ctypes.c_double), R(depth), R(0), R(n - 1), R(0),
      R(n - 1), R(0), R(n - 1), R(gw), R(gw), R(gw), P(u, ctypes.c_double), P(idx, ci), P(sh, ctypes.c_double),
      R(7), P(X, ctypes.c_double), P(V, ctypes.c_double))
    import oracle
    Vo = oracle.interp("IB_4", [h] * 3, [0.0] * 3, [0] * 3, [n - 1] * 3, [gw] * 3, u, X)
    assert np.array_equal(V, Vo)
