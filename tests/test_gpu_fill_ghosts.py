"""Same-level ghost fill (ibamr_plan_fill_ghosts) against a direct evaluation of
its definition: every ghost value whose index, wrapped across periodic
directions, lies in some patch's interior (side box for side data) takes that
patch's value; other ghost values and all interior values are untouched."""
import itertools

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu


def _reference_fill(level, boxes, gcw, side, depth, arrays):
    ndim = level.ndim
    narr = ndim if side else 1
    out = [a.copy() for a in arrays]
    N = level.periodic_shift
    dlo, dhi = level.domain_lower, level.domain_upper
    for p, (lo, hi) in enumerate(boxes):
        for a in range(narr):
            axis = a if side else -1
            top = [hi[d] + (1 if d == axis else 0) for d in range(ndim)]
            rng = [range(lo[d] - gcw[d], top[d] + gcw[d] + 1) for d in range(ndim)]
            for ijk in itertools.product(*rng):
                if all(lo[d] <= ijk[d] <= top[d] for d in range(ndim)):
                    continue
                w, skip = [], False
                for d in range(ndim):
                    if N[d] > 0:
                        w.append(dlo[d] + (ijk[d] - dlo[d]) % N[d])
                    else:
                        if ijk[d] < dlo[d] or ijk[d] > dhi[d] + (1 if d == axis else 0):
                            skip = True
                        w.append(ijk[d])
                if skip:
                    continue
                q = next((q for q, (ql, qh) in enumerate(boxes)
                          if all(ql[d] <= w[d] <= qh[d] + (1 if d == axis else 0) for d in range(ndim))), None)
                if q is None:
                    continue
                qlo = boxes[q][0]
                src = tuple(w[d] - (qlo[d] - gcw[d]) for d in reversed(range(ndim)))
                dst = tuple(ijk[d] - (lo[d] - gcw[d]) for d in reversed(range(ndim)))
                if side:
                    out[p * narr + a][dst] = arrays[q * narr + a][src]
                else:
                    out[p][(slice(None),) + dst] = arrays[q][(slice(None),) + src]
    return out


def _level(ndim, n, boxes, periodic):
    h = 1.0 / n
    patches = [ia.Patch(lo, hi, [l * h for l in lo], [(u + 1) * h for u in hi]) for lo, hi in boxes]
    return ia.Level([h] * ndim, [0] * ndim, [n - 1] * ndim, [0.0] * ndim, [1.0] * ndim,
                    [n if periodic else 0] * ndim, patches)


CASES = [
    # ndim, n, boxes, periodic, side, depth, gcw
    (2, 8, [([0, 0], [3, 3]), ([4, 0], [7, 3]), ([0, 4], [3, 7]), ([4, 4], [7, 7])], True, True, 1, 2),
    (2, 16, [([0, 0], [7, 15]), ([8, 0], [15, 7])], True, False, 2, 3),   # a hole: coarse-fine ghosts stay
    (3, 8, [(list(c), [v + 3 for v in c]) for c in itertools.product([0, 4], repeat=3)], False, True, 1, 2),
    (3, 12, [([2, 2, 2], [5, 9, 5]), ([6, 2, 2], [9, 5, 9])], False, False, 1, 3),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dd-n%d-%dp-%s-%s" % (
    c[0], c[1], len(c[2]), "per" if c[3] else "nonper", "side" if c[4] else "cell"))
def test_fill_ghosts_matches_definition(case):
    ndim, n, boxes, periodic, side, depth, g = case
    level = _level(ndim, n, boxes, periodic)
    cent = ia.SIDE if side else ia.CELL
    gcw = [g] * ndim
    plan = ia.LevelPlan(level, "PIECEWISE_LINEAR", cent, ia.FP64, depth=depth, data_gcw=gcw, max_points=1)
    rng = np.random.default_rng(7)
    arrays = []
    for p in range(len(boxes)):
        for a in range(plan.narr):
            arrays.append(rng.uniform(-1, 1, ia.array_shape(level, p, gcw, cent, axis=a, depth=depth)))
    dev = [torch.from_numpy(a.copy()).cuda() for a in arrays]
    plan.bind(0, dev)
    print("fill ghosts", ndim, n, len(boxes), flush=True)
    plan.fill_ghosts(0)
    torch.cuda.synchronize()
    ref = _reference_fill(level, boxes, gcw, side, depth, arrays)
    for i, (d, r) in enumerate(zip(dev, ref)):
        assert np.array_equal(d.cpu().numpy(), r), i


def test_fill_ghosts_equals_periodic_fill_for_consistent_data():
    n, g = 16, 3
    level = ia.Level.uniform_periodic(n, 3)
    plan = ia.LevelPlan(level, "IB_4", ia.SIDE, ia.FP64, data_gcw=[g] * 3, max_points=1)
    rng = np.random.default_rng(8)
    arrays = []
    for a in range(3):
        u = rng.uniform(-1, 1, ia.array_shape(level, 0, [g] * 3, ia.SIDE, axis=a))
        idx = [slice(None)] * 3
        ax = 2 - a
        src = list(idx)
        src[ax] = g
        dst = list(idx)
        dst[ax] = g + n
        u[tuple(dst)] = u[tuple(src)]   # face N equals face 0 (one degree of freedom)
        arrays.append(u)
    d1 = [torch.from_numpy(a.copy()).cuda() for a in arrays]
    d2 = [torch.from_numpy(a.copy()).cuda() for a in arrays]
    plan.bind(0, d1)
    plan.bind(1, d2)
    plan.fill_ghosts(0)
    plan.fill_periodic_ghosts(1)
    torch.cuda.synchronize()
    for a in range(3):
        assert torch.equal(d1[a], d2[a]), a
