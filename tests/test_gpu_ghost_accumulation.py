"""GPU level plans against the reference's multi-patch ghost-accumulation
outputs (tests/IBTK/ghost_accumulation_01_*.spread.*.output; PIECEWISE_LINEAR,
cell and side data, one to eight patches, periodic and non-periodic): one FP64
spread launch over all patches of the finest level must equal the oracle
bitwise and the reference's accumulated values to summation order (1e-13)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
import ghost_cases as gc  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", gc.CASES)
def test_level_spread_matches_reference_accumulation(orc, case):
    c = gc.load(case)
    ndim, h, nf = c["ndim"], c["h"], c["nf"]
    keep = gc.level_points(orc, c)
    X, Q = np.ascontiguousarray(c["X"][keep]), np.ascontiguousarray(c["Q"][keep])
    patches = [ia.Patch(lo, hi, xl, xu) for lo, hi, xl, xu in c["boxes"]]
    level = ia.Level([h] * ndim, [0] * ndim, [nf - 1] * ndim, [0.0] * ndim, [1.0] * ndim, c["pshift"], patches)
    cent = ia.SIDE if c["side"] else ia.CELL
    depth = 1 if c["side"] else Q.shape[1]
    plan = ia.LevelPlan(level, c["kernel"], cent, ia.FP64, depth=depth, data_gcw=c["gcw"], idx_gcw=c["gcw"][0],
                        max_points=max(1, X.shape[0]))
    arrays = []
    for p in range(len(patches)):
        for a in range(plan.narr):
            arrays.append(torch.zeros(ia.array_shape(level, p, c["gcw"], cent, axis=a, depth=depth),
                                      dtype=torch.float64, device="cuda"))
    print("ghost accumulation", case, len(patches), "patches", X.shape[0], "points", flush=True)
    Xd = torch.from_numpy(X).cuda()
    plan.bin(Xd)
    plan.bind(0, arrays)
    plan.spread(0, torch.from_numpy(Q).cuda(), Xd, mode=ia.ACCUMULATE)
    torch.cuda.synchronize()
    cells = orc.get_cell_index(X, [0.0] * ndim, [1.0] * ndim, [h] * ndim, [0] * ndim, [nf - 1] * ndim)
    n = 0
    for pi, (lo, hi, xl, xu) in enumerate(c["boxes"]):
        got = [arrays[pi * plan.narr + a].cpu().numpy() for a in range(plan.narr)]
        if c["side"]:
            ref = [np.zeros(orc.side_axis_shape(lo, hi, c["gcw"], a)) for a in range(ndim)]
        else:
            ref = np.zeros(orc.ghosted_shape(lo, hi, c["gcw"], depth=depth))
        orc.spread_patch(c["kernel"], c["side"], [h] * ndim, xl, lo, hi, c["gcw"], ref, X, Q, cells, c["gcw"][0],
                         c["pshift"], [0] * ndim, depth=depth)
        if c["side"]:
            for a in range(ndim):
                assert np.array_equal(got[a], ref[a]), (pi, a)
        else:
            assert np.array_equal(got[0], ref), pi
        n += gc.compare(c, pi, got)
    assert n > 0
