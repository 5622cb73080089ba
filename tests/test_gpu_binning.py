"""Binning pinned directly: the cell index the GPU binning computes for every
point (ibamr_plan_get_cell_indices) equals the oracle's getCellIndex
(IndexUtilities-inl.h:52-72, level form :226-243) bit for bit:

  * the full C2 workload (256^3, 2^20 points on a jittered Fibonacci sphere);
  * points exactly on cell faces and one ulp either side, at x_lower, at
    x_upper (which takes the upper branch and lands one past the box) and past
    the middle of the domain (the x_upper branch), on a level whose index space
    starts at -5 with a spacing that is not a power of two;
  * positions that are NaN or too far away for an int give INT_MIN, and are
    reported outside the domain.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu
INT_MIN = -(1 << 31)


def _cells(level, X, kernel="IB_4"):
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=X.shape[0])
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to("cuda")
    plan.bin(Xd, check=False)
    c = plan.cell_indices(X.shape[0]).cpu().numpy()
    return plan, c


def test_config2_cells_bitwise(orc):
    n, N = 1 << 20, 256
    h = 1.0 / N
    rng = np.random.default_rng(202)
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    r = np.sqrt(1 - z * z)
    phi = math.pi * (3 - math.sqrt(5)) * np.arange(n)
    X = 0.5 + 0.3 * np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    X += rng.uniform(-h / 4, h / 4, X.shape)
    print("binning C2", flush=True)
    _, got = _cells(ia.Level.uniform_periodic(N, 3), X)
    ref = orc.get_cell_index(X, [0.0] * 3, [1.0] * 3, [h] * 3, [0] * 3, [N - 1] * 3)
    assert np.array_equal(got, ref)


def _offset_level(ndim, lo=-5, n=64, dx=0.1, xlo=-0.5):
    p = ia.Patch([lo] * ndim, [lo + n - 1] * ndim, [xlo] * ndim, [xlo + n * dx] * ndim)
    return ia.Level([dx] * ndim, [lo] * ndim, [lo + n - 1] * ndim, [xlo] * ndim, [xlo + n * dx] * ndim,
                    [0] * ndim, [p])


@pytest.mark.parametrize("ndim", [2, 3])
def test_faces_ulps_and_upper_branch(orc, ndim):
    lo, n, dx, xlo = -5, 64, 0.1, -0.5
    xhi = xlo + n * dx
    faces = xlo + dx * np.arange(n + 1)
    vals = np.concatenate([faces, np.nextafter(faces, -np.inf), np.nextafter(faces, np.inf),
                           xlo + dx * (np.arange(n) + 0.5), [xlo, xhi, 0.5 * (xlo + xhi),
                                                             np.nextafter(0.5 * (xlo + xhi), np.inf)]])
    rng = np.random.default_rng(31)
    X = np.stack([rng.permutation(vals) for _ in range(ndim)], axis=1)
    print("binning faces", ndim, X.shape[0], flush=True)
    plan, got = _cells(_offset_level(ndim, lo, n, dx, xlo), X)
    ref = orc.get_cell_index(X, [xlo] * ndim, [xhi] * ndim, [dx] * ndim, [lo] * ndim, [lo + n - 1] * ndim)
    assert np.array_equal(got, ref)
    # x_upper itself: the upper branch gives iupper + 1 (outside the box)
    k = int(np.nonzero(X[:, 0] == xhi)[0][0])
    assert got[k, 0] == lo + n
    with pytest.raises(ia.IBAMRError) as e:
        plan.check()
    assert e.value.code == ia.IBAMR_ERR_OUT_OF_DOMAIN


def test_nan_and_far_positions_give_int_min(orc):
    level = _offset_level(3)
    X = np.array([[0.1, 0.2, 0.3], [np.nan, 0.2, 0.3], [0.1, 1e300, 0.3], [0.1, 0.2, -1e12]])
    print("binning nan/far", flush=True)
    plan, got = _cells(level, X)
    ref = orc.get_cell_index(X[:1], [-0.5] * 3, [5.9] * 3, [0.1] * 3, [-5] * 3, [58] * 3)
    assert np.array_equal(got[0], ref[0])
    assert got[1, 0] == INT_MIN and got[2, 1] == INT_MIN and got[3, 2] == INT_MIN
    assert np.array_equal(got[1, 1:], got[0, 1:])
    with pytest.raises(ia.IBAMRError) as e:
        plan.check()
    assert e.value.code == ia.IBAMR_ERR_OUT_OF_DOMAIN
    assert "3 point(s)" in str(e.value)
