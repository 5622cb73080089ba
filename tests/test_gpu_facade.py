"""GPU tests of the C++ layer (include/ibtk_amd.hpp) and of the single-patch
raw-array plans (ibamr_plan_create_patch) behind it.

  * tests/cpp/interpolate_01 interp <kernel>: the reference's
    tests/interpolate/interpolate_01.cpp (3-D) on ibtk_amd::LEInteractor; its
    output must match interpolate_01_3d.<kernel>.output (golden fixture) to the
    16 printed digits.
  * ... side <kernel>: raw-array side interp over a grown box and spread over the
    patch box, against the oracle (interp bitwise; spread to 1e-13 relative:
    the summation order differs from the oracle's list order).
  * ... level <kernel>: LDataManager / IBMethod on a periodic level: f holds 1
    before the spread and keeps 1 in its ghost cells; U after the same-level
    ghost fill; X_new = X + dt U bitwise.
  * ... regrid <kernel>: IBStandardForceGen then LDataManager::redistribute on a
    two-patch periodic level: forces bitwise with the oracle at the old
    positions, positions / order / Lagrangian indices / patch offsets exactly
    as the oracle's redistribution, and F, U carried to the new order.
  * LevelPlan(box=...) in Python for cell and side data and several boxes.
"""
import json
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "bin", "interpolate_01")
GOLD = os.path.join(ROOT, "tests", "golden")
KERNELS = [("IB_4", "ib_4"), ("IB_6", "ib_6"), ("BSPLINE_3", "bspline_3"), ("BSPLINE_4", "bspline_4"),
           ("BSPLINE_5", "bspline_5"), ("BSPLINE_6", "bspline_6")]
# print precision; BSPLINE_4/5/6: the conditioning of their polynomial form (test_oracle_golden.COND_TOL)
TOL = {"BSPLINE_4": 5e-14, "BSPLINE_5": 5e-13, "BSPLINE_6": 3e-12}


def _run(mode, kernel, tmp_path):
    assert os.path.exists(DRIVER), "C++ driver not built (python -m ibamr_amd.build)"
    out = tmp_path / ("%s_%s.txt" % (mode, kernel))
    print("driver", mode, kernel, flush=True)
    r = subprocess.run([DRIVER, mode, kernel, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return out.read_text().splitlines()


@pytest.mark.parametrize("kernel,tag", KERNELS)
def test_cpp_interpolate_01_3d_matches_reference_output(kernel, tag, tmp_path):
    lines = _run("interp", kernel, tmp_path)
    assert lines[0] == "x, y, z, u0, u1, u2, e0, e1, e2"
    got = np.array([[float(v) for v in l.split(",")] for l in lines[1:]])
    gold = np.array(json.load(open(os.path.join(GOLD, "interpolate_01_3d.%s.json" % tag)))["rows"])
    assert got.shape == gold.shape == (100, 6)
    assert np.max(np.abs(got - gold) / np.abs(gold)) < TOL.get(kernel, 2e-15)


def _parse_blocks(lines, npts, ncol):
    rows = np.array([[float(v) for v in l.split()] for l in lines[1:1 + npts]])
    assert rows.shape == (npts, ncol)
    blocks = {}
    for l in lines[1 + npts:]:
        name, *vals = l.split()
        blocks[name] = np.array([float(v) for v in vals])
    return rows, blocks


@pytest.mark.parametrize("kernel", ["IB_4", "IB_6", "BSPLINE_3"])
def test_cpp_side_raw_array(kernel, tmp_path, orc):
    rows, blocks = _parse_blocks(_run("side", kernel, tmp_path), 200, 9)
    X, F, U = rows[:, :3], rows[:, 3:6], rows[:, 6:9]
    gw = orc.min_ghost_width(kernel)
    il, iu, dx, xl, xu = [8] * 3, [15] * 3, [1 / 32] * 3, [0.25] * 3, [0.5] * 3
    u = [blocks["u%d" % a].reshape(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    lo, hi = [7] * 3, [16] * 3
    idx = orc.build_local_indices(lo, hi, xl, xu, dx, il, iu, X)
    assert 0 < len(idx) < 200
    ref = orc.interp_side(kernel, dx, xl, il, iu, [gw] * 3, u, X, indices=idx)
    taken = np.zeros(200, dtype=bool)
    taken[idx] = True
    assert np.array_equal(U[taken], ref[taken])           # bitwise
    assert np.all(U[~taken] == -7.0)                      # untouched
    # spread over the patch box: the oracle over the same selection
    f = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    sidx = orc.build_local_indices(il, iu, xl, xu, dx, il, iu, X)
    orc.spread_side(kernel, dx, xl, il, iu, [gw] * 3, f, X, F, indices=sidx)
    for a in range(3):
        got = blocks["f%d" % a]
        want = f[a].reshape(-1)
        assert got.shape == want.shape
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("kernel", ["IB_4", "BSPLINE_3"])
def test_cpp_ldatamanager_ibmethod_level(kernel, tmp_path, orc):
    N = 16
    rows, blocks = _parse_blocks(_run("level", kernel, tmp_path), 300, 12)
    X, F, U, Xn = rows[:, :3], rows[:, 3:6], rows[:, 6:9], rows[:, 9:12]
    gw = orc.min_ghost_width(kernel)
    h = 1.0 / N
    il, iu = [0] * 3, [N - 1] * 3
    assert np.array_equal(Xn, X + 0.01 * U)
    # u after the same-level ghost fill: every ghost value equals its periodic
    # image; the interior face N of the side axis is interior data and keeps its value
    u_fill = []
    for a in range(3):
        v = blocks["u%d" % a].reshape(orc.side_axis_shape(il, iu, [gw] * 3, a))
        w = v
        for ax in range(3):
            w = np.take(w, (np.arange(w.shape[ax]) - gw) % N + gw, axis=ax)
        mask = np.ones(v.shape, dtype=bool)
        sl = [slice(gw, gw + N)] * 3
        sl[2 - a] = gw + N
        mask[tuple(sl)] = False
        assert np.array_equal(v[mask], w[mask]), a
        u_fill.append(v)
    cells = orc.get_cell_index(X, [0.0] * 3, [1.0] * 3, [h] * 3, il, iu)
    iidx, ish = orc.build_patch_entries(cells, il, iu, 0, [N] * 3, il, [h] * 3, [8] * 3, interior_only=True)
    ref = orc.interp_side(kernel, [h] * 3, [0.0] * 3, il, iu, [gw] * 3, u_fill, X, indices=iidx, Xshift=ish)
    assert np.array_equal(U, ref)
    # spread: interior = 1 + S F (all periodic images), ghosts stay 1
    f = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    sidx, ssh = orc.build_patch_entries(cells, il, iu, gw, [N] * 3, il, [h] * 3, [8] * 3)
    orc.spread_side(kernel, [h] * 3, [0.0] * 3, il, iu, [gw] * 3, f, X, F, indices=sidx, Xshift=ssh)
    for a in range(3):
        got = blocks["f%d" % a].reshape(f[a].shape)
        inner = np.zeros(f[a].shape, dtype=bool)
        sl = [slice(gw, gw + N + (1 if (2 - ax) == a else 0)) for ax in range(3)]
        inner[tuple(sl)] = True
        assert np.all(got[~inner] == 1.0)
        want = 1.0 + f[a][inner]
        assert np.max(np.abs(got[inner] - want)) <= 1e-13 * np.max(np.abs(want))


# ----------------------------------------------------------------------------- raw-array plans from Python

def _patch_level(ndim, lo, n, h, xlo):
    p = ia.Patch([lo] * ndim, [lo + n - 1] * ndim, [xlo] * ndim, [xlo + n * h] * ndim)
    return ia.Level([h] * ndim, [lo] * ndim, [lo + n - 1] * ndim, [xlo] * ndim, [xlo + n * h] * ndim, [0] * ndim, [p])


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("centering", [ia.CELL, ia.SIDE])
@pytest.mark.parametrize("grow", [0, 1, -2])
def test_patch_plan_interp_bitwise(orc, ndim, centering, grow):
    kernel, lo, n, h, xlo = "IB_4", 4, 12, 1.0 / 40, 0.1
    level = _patch_level(ndim, lo, n, h, xlo)
    gw = orc.min_ghost_width(kernel)
    box = ([lo - grow] * ndim, [lo + n - 1 + grow] * ndim)
    rng = np.random.default_rng(100 + ndim + 10 * grow)
    X = rng.uniform(xlo - 2.5 * h, xlo + (n + 2.5) * h, (500, ndim))
    depth = 2 if centering == ia.CELL else 1
    plan = ia.LevelPlan(level, kernel, centering, ia.FP64, depth=depth, max_points=500, box=box)
    arrs = []
    for a in range(plan.narr):
        shp = ia.array_shape(level, 0, plan.data_gcw, centering, axis=a, depth=depth)
        arrs.append(rng.uniform(-1, 1, shp))
    td = [torch.tensor(a, device="cuda") for a in arrs]
    plan.bind(0, td)
    Xd = torch.tensor(X, device="cuda")
    print("patch plan interp", ndim, centering, grow, flush=True)
    plan.bin(Xd, check=False)
    nq = ndim if centering == ia.SIDE else depth
    U = torch.full((500, nq), -3.0, dtype=torch.float64, device="cuda")
    plan.interp(0, U, Xd)
    U = U.cpu().numpy()
    il, iu = [lo] * ndim, [lo + n - 1] * ndim
    idx = orc.build_local_indices(box[0], box[1], [xlo] * ndim, [xlo + n * h] * ndim, [h] * ndim, il, iu, X)
    taken = np.zeros(500, dtype=bool)
    taken[idx] = True
    if centering == ia.SIDE:
        ref = orc.interp_side(kernel, [h] * ndim, [xlo] * ndim, il, iu, [gw] * ndim, arrs, X, indices=idx)
    else:
        ref = orc.interp(kernel, [h] * ndim, [xlo] * ndim, il, iu, [gw] * ndim, arrs[0], X, indices=idx, depth=depth)
    assert np.array_equal(U[taken], ref[taken])
    assert np.all(U[~taken] == -3.0)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("grow", [0, 2])
def test_patch_plan_spread_and_interior_mode(orc, ndim, grow):
    kernel, lo, n, h, xlo = "IB_6", 0, 16, 1.0 / 16, 0.0
    level = _patch_level(ndim, lo, n, h, xlo)
    gw = orc.min_ghost_width(kernel)
    box = ([lo - grow] * ndim, [lo + n - 1 + grow] * ndim)
    rng = np.random.default_rng(11 + ndim)
    X = rng.uniform(-3 * h, 1 + 3 * h, (700, ndim))
    F = rng.standard_normal((700, ndim))
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=700, box=box)
    base = [rng.uniform(-1, 1, ia.array_shape(level, 0, plan.data_gcw, ia.SIDE, axis=a)) for a in range(ndim)]
    Xd, Fd = torch.tensor(X, device="cuda"), torch.tensor(F, device="cuda")
    print("patch plan spread", ndim, grow, flush=True)
    plan.bin(Xd, check=False)
    il, iu = [lo] * ndim, [lo + n - 1] * ndim
    idx = orc.build_local_indices(box[0], box[1], [xlo] * ndim, [1.0] * ndim, [h] * ndim, il, iu, X)
    ref = [b.copy() for b in base]
    orc.spread_side(kernel, [h] * ndim, [xlo] * ndim, il, iu, [gw] * ndim, ref, X, F, indices=idx)
    for mode in (ia.ACCUMULATE, ia.ACCUMULATE_INTERIOR):
        td = [torch.tensor(b, device="cuda") for b in base]
        plan.bind(0, td)
        plan.spread(0, Fd, Xd, mode=mode)
        torch.cuda.synchronize()
        for a in range(ndim):
            got = td[a].cpu().numpy()
            scale = np.max(np.abs(ref[a]))
            if mode == ia.ACCUMULATE:
                assert np.max(np.abs(got - ref[a])) <= 1e-13 * scale
            else:
                inner = np.zeros(got.shape, dtype=bool)
                sl = [slice(gw, gw + n + (1 if (ndim - 1 - ax) == a else 0)) for ax in range(ndim)]
                inner[tuple(sl)] = True
                assert np.array_equal(got[~inner], base[a][~inner])
                assert np.max(np.abs(got[inner] - ref[a][inner])) <= 1e-13 * scale


def test_cpp_forcegen_and_redistribute(tmp_path, orc):
    lines = _run("regrid", "IB_4", tmp_path)
    n = 400
    rows = np.array([[float(v) for v in l.split()] for l in lines[1:1 + n]])
    Xb, Ub, Xa, Fa, Ua, la = rows[:, 0:3], rows[:, 3:6], rows[:, 6:9], rows[:, 9:12], rows[:, 12:15], rows[:, 15]
    offsets = [int(v) for v in lines[1 + n].split()[1:]]
    i = np.arange(n)
    springs = (i, (i + 1) % n, 1.0 + 0.01 * i, np.full(n, 0.05))
    beams = ((i - 1) % n, i, (i + 1) % n, np.full(n, 0.5), np.tile([0.001, -0.002, 0.0], (n, 1)))
    tn = np.arange(0, n, 7)
    targets = (tn, np.full(len(tn), 3.0), np.full(len(tn), 0.5), np.full((len(tn), 3), 0.5))
    # spring i under force function i % 3, registered as Hookean, linear strain, polynomial {k, 0.05, 30, -40}
    law = (i % 3).astype(np.int32)
    par = [[1.0 + 0.01 * k, 0.05] + ([30.0, -40.0] if k % 3 == 2 else []) for k in range(n)]
    off = np.concatenate([[0], np.cumsum([len(p) for p in par])]).astype(np.int32)
    F = orc.lagrangian_force(3, Xb, U=Ub, springs=springs, beams=beams, targets=targets,
                             spring_laws=(law, off, np.array([v for p in par for v in p])))
    lag = n - 1 - i
    boxes = [([0, 0, 0], [7, 15, 15]), ([8, 0, 0], [15, 15, 15])]
    rc, Xw, order, poff, _ = orc.redistribute(Xb, [0.0] * 3, [1.0] * 3, [1.0 / 16] * 3, [0] * 3, [15] * 3,
                                              [True] * 3, boxes, lag=lag)
    assert rc == 0
    assert np.array_equal(Xa, Xw[order])
    assert np.array_equal(Fa, F[order])
    assert np.array_equal(Ua, Ub[order])
    assert np.array_equal(la, lag[order])
    assert offsets == list(poff)
