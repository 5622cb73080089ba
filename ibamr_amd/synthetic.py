"""Seeded synthetic workloads (host side; numpy): point clouds and an AMR level.

amr_level(): the fine level of a two-level hierarchy, coarse n^3 cells on the
unit cube refined by `ratio`.  Points lie on `nshell` ellipsoidal shells
(centres uniform in [0.25, 0.75]^3, semi-axes uniform in [0.05, 0.15]).  The
fine index space is cut into bricks without overlap or gaps: the z axis into
slabs, each slab's y axis into rows, each row's x axis into bricks, every cut
length drawn uniformly and independently from {16, 24, 32, 48, 64} (scaled by
block / 64), so each brick's three edges are uniform draws from that set.  The
bricks that hold points become patches -- every point lies in exactly one
patch interior, as LDataManager requires of a fine level -- and, when
`npatches` asks for more, the bricks nearest the shells are added (the buffer
of refined cells an AMR tagger keeps around a structure) until the level has
`npatches` patches.
"""
import numpy as np

import ibamr_amd as ia

EDGES = (16, 24, 32, 48, 64)


def shell_points(nshell, per_shell, rng, centre=(0.25, 0.75), axes=(0.05, 0.15)):
    X = []
    for _ in range(nshell):
        c = rng.uniform(centre[0], centre[1], 3)
        ax = rng.uniform(axes[0], axes[1], 3)
        d = rng.standard_normal((per_shell, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        X.append(c + d * ax)
    return np.ascontiguousarray(np.concatenate(X))


def _cuts(rng, n, edges):
    """0 = c_0 < c_1 < ... < c_m = n with every c_{k+1} - c_k drawn uniformly
    from `edges` (a sequence of draws that passes n is drawn again)."""
    while True:
        c = [0]
        while c[-1] < n:
            c.append(c[-1] + int(edges[rng.integers(len(edges))]))
        if c[-1] == n:
            return np.array(c, dtype=np.int64)


def _lattice_distance(mask):
    """City-block distance (in lattice cells) of every cell of a 3-D boolean
    lattice to the nearest True cell (non-periodic); -1 where none is reachable."""
    dist = np.where(mask, 0, -1).astype(np.int64)
    cur = mask.copy()
    step = 0
    while True:
        step += 1
        nxt = cur.copy()
        for ax in range(3):
            for sh in (1, -1):
                r = np.roll(cur, sh, axis=ax)
                edge = [slice(None)] * 3
                edge[ax] = 0 if sh == 1 else -1
                r[tuple(edge)] = False
                nxt |= r
        new = nxt & ~cur
        if not new.any():
            return dist
        dist[new] = step
        cur = nxt


def amr_level(coarse=128, ratio=4, npoints=1 << 21, nshell=4, seed=2028, block=64, npatches=None):
    """Returns (level, X): an ibamr_amd.Level of fine patches and float64 points [n][3].

    npatches: the number of patches (occupied bricks, then the nearest empty
    ones); None keeps the occupied bricks only.  Raises ValueError when the
    points occupy more bricks than npatches."""
    rng = np.random.default_rng(seed)
    N = coarse * ratio
    h = 1.0 / N
    edges = [e * block // 64 for e in EDGES]
    unit = int(np.gcd.reduce(edges))
    if N % unit:
        raise ValueError("fine grid %d is not a multiple of the edge lattice %d" % (N, unit))
    X = shell_points(nshell, npoints // nshell, rng)
    cell = np.floor(X / h).astype(np.int64)
    if cell.min() < 0 or cell.max() >= N:
        raise ValueError("points outside the unit cube")
    zc = _cuts(rng, N, edges)
    yc = [_cuts(rng, N, edges) for _ in range(len(zc) - 1)]
    xc = [[_cuts(rng, N, edges) for _ in range(len(y) - 1)] for y in yc]
    bricks = [(k, j, i) for k in range(len(zc) - 1) for j in range(len(yc[k]) - 1) for i in range(len(xc[k][j]) - 1)]

    def box(b):
        k, j, i = b
        lo = [int(xc[k][j][i]), int(yc[k][j]), int(zc[k])]
        hi = [int(xc[k][j][i + 1]) - 1, int(yc[k][j + 1]) - 1, int(zc[k + 1]) - 1]
        return lo, hi

    zi = np.searchsorted(zc, cell[:, 2], side="right") - 1
    occupied = set()
    for k in np.unique(zi):
        sel = cell[zi == k]
        yi = np.searchsorted(yc[k], sel[:, 1], side="right") - 1
        for j in np.unique(yi):
            xi = np.searchsorted(xc[k][j], sel[yi == j][:, 0], side="right") - 1
            occupied.update((int(k), int(j), int(i)) for i in np.unique(xi))
    chosen = sorted(occupied)
    if npatches is not None:
        if len(chosen) > npatches:
            raise ValueError("the points occupy %d bricks, more than %d patches" % (len(chosen), npatches))
        G = N // unit
        mask = np.zeros((G, G, G), dtype=bool)
        mask[cell[:, 0] // unit, cell[:, 1] // unit, cell[:, 2] // unit] = True
        dist = _lattice_distance(mask)
        cand = []
        for b in bricks:
            if b in occupied:
                continue
            lo, hi = box(b)
            d = dist[lo[0] // unit:(hi[0] + 1) // unit, lo[1] // unit:(hi[1] + 1) // unit,
                     lo[2] // unit:(hi[2] + 1) // unit]
            d = d[d >= 0]
            if d.size:
                cand.append((int(d.min()), b))
        cand.sort()
        chosen = sorted(chosen + [b for _, b in cand[:npatches - len(chosen)]])
        if len(chosen) < npatches:
            raise ValueError("only %d bricks in the domain, fewer than %d patches" % (len(chosen), npatches))
    patches = []
    for b in chosen:
        lo, hi = box(b)
        patches.append(ia.Patch(lo, hi, [i * h for i in lo], [(i + 1) * h for i in hi]))
    level = ia.Level([h] * 3, [0] * 3, [N - 1] * 3, [0.0] * 3, [1.0] * 3, [0] * 3, patches)
    return level, X
