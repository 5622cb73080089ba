"""Set-up of the reference's tests/IBTK/ghost_accumulation_01 spread cases
(fixtures tests/golden/ghost_accumulation_01_<case>.patches.json): the finest
level's patches, the 100 points and values drawn with std::mt19937(42) (Q
first, then X, ghost_accumulation_01.cpp:131-135), and the comparison of a
computed field with the accumulated values the reference printed."""
import json
import os

import numpy as np

from mt19937 import MT19937, uniform_real

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["2d.cell.spread.a", "2d.cell.spread.b", "2d.side.spread.a", "2d.side.spread.b", "3d.cell.spread.a",
         "3d.cell.spread.b", "3d.side.spread.a", "3d.side.spread.b"]


def load(case):
    with open(os.path.join(GOLD, "ghost_accumulation_01_%s.patches.json" % case)) as f:
        fx = json.load(f)
    p = fx["params"]
    ndim = fx["ndim"]
    nf = p["N"] * p["ratio"]
    h = 1.0 / nf
    boxes = []
    for pt in fx["patches"]:
        lo = [int(round(v / h)) for v in pt["x_lower"]]
        hi = [int(round(v / h)) - 1 for v in pt["x_upper"]]
        boxes.append((lo, hi, pt["x_lower"], pt["x_upper"]))
    side = p["var_type"] == "SIDE"
    qdepth = ndim if side else p["depth"]
    g = MT19937(42)
    n = 100
    Q = np.array([uniform_real(g, 0.0, 1.0) for _ in range(qdepth * n)]).reshape(n, qdepth)
    X = np.array([uniform_real(g, 0.0, 1.0) for _ in range(ndim * n)]).reshape(n, ndim)
    return {"fx": fx, "ndim": ndim, "nf": nf, "h": h, "boxes": boxes, "side": side, "Q": Q, "X": X,
            "gcw": [p["ghost_width"]] * ndim, "kernel": p["kernel"],
            "pshift": [nf * v for v in p["periodic"]]}


def level_points(orc, c):
    """Points on the finest level: those whose cell lies in one of its patches
    (all of them when the level covers the domain)."""
    ndim, nf, h = c["ndim"], c["nf"], c["h"]
    cells = orc.get_cell_index(c["X"], [0.0] * ndim, [1.0] * ndim, [h] * ndim, [0] * ndim, [nf - 1] * ndim)
    keep = np.zeros(len(cells), dtype=bool)
    for lo, hi, _, _ in c["boxes"]:
        keep |= np.all((cells >= np.array(lo)) & (cells <= np.array(hi)), axis=1)
    return keep


def compare(c, patch_index, arrays, tol=1e-13):
    """arrays: the patch's ghosted arrays (side: one per axis, numpy order
    (z,)y,x; cell: one (depth, ...) array).  Returns the number of values checked."""
    lo, hi, _, _ = c["boxes"][patch_index]
    ndim, gcw = c["ndim"], c["gcw"]
    ref = c["fx"]["patches"][patch_index]["arrays"]
    scale = max(abs(e[-1]) for pt in c["fx"]["patches"] for k in pt["arrays"] for e in pt["arrays"][k])
    n = 0
    for key, entries in ref.items():
        a = int(key)
        arr = arrays[a] if c["side"] else arrays[0][a]
        for e in entries:
            idx = tuple(e[ndim - 1 - d] - (lo[ndim - 1 - d] - gcw[ndim - 1 - d]) for d in range(ndim))
            got = float(arr[idx])
            v = e[-1]
            assert abs(got - v) <= tol * scale, (patch_index, key, e, got)
            n += 1
    return n
