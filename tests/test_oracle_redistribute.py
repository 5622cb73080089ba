"""The oracle's restatement of LDataManager's redistribution (wrap/clamp into
the domain, periodic offsets, owner patch, new local order) pinned by closed
forms of the wrap and clamp arithmetic (LDataManager.cpp:1393-1431), and by a
direct evaluation of the order's definition: patches in level order, cells in
box order, Lagrangian index (numpy lexsort over independently computed keys)."""
import numpy as np
import pytest

H = 1.0 / 16


def _grid2(periodic=(True, True)):
    boxes = [((0, 0), (7, 15)), ((8, 0), (15, 7)), ((8, 8), (15, 15))]
    return dict(x_lower=[0.0, 0.0], x_upper=[1.0, 1.0], dx=[H, H], dom_lo=[0, 0], dom_hi=[15, 15],
                periodic=periodic, boxes=boxes)


def test_periodic_wrap_closed_forms(orc):
    g = _grid2()
    X = np.array([[1.0, 0.5], [-1e-17, 0.5], [3.25, -2.75], [0.5, 0.999999]])
    rc, Xn, order, poff, offs = orc.redistribute(X, **g)
    assert rc == 0
    # x = 1.0 -> 0.0 (one length); -1e-17 + 1 rounds to 1.0, then wraps to 0.0
    assert Xn[0, 0] == 0.0 and offs[0, 0] == 16
    assert Xn[1, 0] == 0.0 and offs[1, 0] == 0
    assert Xn[2, 0] == 0.25 and offs[2, 0] == 48 and Xn[2, 1] == 0.25 and offs[2, 1] == -48
    assert Xn[3, 1] == 0.999999 and offs[3, 1] == 0


def test_clamp_and_escape(orc):
    g = _grid2(periodic=(False, False))
    X = np.array([[-0.1, 0.5], [1.5, 0.25], [0.3, 1.0]])
    rc, Xn, *_ = orc.redistribute(X, **g)
    assert rc == 0
    assert Xn[0, 0] == 0.0 and Xn[1, 0] == 1.0 - np.sqrt(np.finfo(float).eps)
    assert Xn[2, 1] == 1.0 - np.sqrt(np.finfo(float).eps)
    rc, *_ = orc.redistribute(X, error_if_leave=True, **g)
    assert rc == 1
    # exactly on x_upper is not an escape, but its cell lies outside the domain: in no patch
    rc, *_ = orc.redistribute(np.array([[1.0, 0.5]]), error_if_leave=True, **g)
    assert rc == 2


def test_no_patch_and_lost(orc):
    g = _grid2()
    g["boxes"] = g["boxes"][:2]  # the upper right quarter is not covered
    rc, *_ = orc.redistribute(np.array([[0.75, 0.75]]), **g)
    assert rc == 2
    rc, *_ = orc.redistribute(np.array([[0.5, 70.0]]), **_grid2())
    assert rc == 3
    rc, *_ = orc.redistribute(np.array([[np.nan, 0.5]]), **_grid2())
    assert rc == 3


@pytest.mark.parametrize("ndim", [2, 3])
def test_order_matches_definition(orc, ndim):
    rng = np.random.default_rng(7 + ndim)
    N = 16
    if ndim == 2:
        boxes = _grid2()["boxes"]
    else:
        boxes = [((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 7)), ((8, 0, 8), (15, 15, 15))]
    n = 3000
    X = rng.uniform(-1.2, 2.2, (n, ndim))
    X[:50] = X[50:100]  # coincident points: ordered by Lagrangian index
    lag = rng.permutation(n) + 5
    rc, Xn, order, poff, offs = orc.redistribute(X, [0.0] * ndim, [1.0] * ndim, [1.0 / N] * ndim, [0] * ndim,
                                                 [N - 1] * ndim, [True] * ndim, boxes, lag=lag)
    assert rc == 0
    want = np.mod(X, 1.0)
    assert np.allclose(Xn, want, atol=1e-15)
    assert np.array_equal(offs, np.round((X - Xn) * N).astype(int))
    cell = np.floor(Xn * N).astype(int)
    patch = np.full(n, -1)
    lin = np.zeros(n, dtype=np.int64)
    for q, (lo, hi) in enumerate(boxes):
        inside = np.all((cell >= lo) & (cell <= hi), axis=1) & (patch < 0)
        patch[inside] = q
        ext = np.array(hi) - np.array(lo) + 1
        rel = cell - np.array(lo)
        l = np.zeros(n, dtype=np.int64)
        for d in reversed(range(ndim)):
            l = l * ext[d] + rel[:, d]
        lin[inside] = l[inside]
    assert np.all(patch >= 0)
    ref = np.lexsort((np.arange(n), lag, lin, patch))
    assert np.array_equal(order, ref)
    assert np.array_equal(poff, np.concatenate([[0], np.cumsum(np.bincount(patch, minlength=len(boxes)))]))
