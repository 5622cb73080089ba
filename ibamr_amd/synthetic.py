"""Seeded synthetic workloads (host side; numpy): point clouds and an AMR level.

amr_level(): the fine level of a two-level hierarchy, coarse n^3 cells on the
unit cube refined by `ratio`.  Points lie on `nshell` ellipsoidal shells
(centres uniform in [0.25, 0.75]^3, semi-axes uniform in [0.05, 0.15]).  The
fine patches tile the occupied part of the fine index space without overlap:
every 64^3 block that holds points is split along each axis into segments
drawn from {16, 24, 32, 48, 64} (a random partition of 64 per axis; a smaller
`block` scales the segments with it), and the
pieces that hold at least one point become patches, so every point lies in
exactly one patch interior (as LDataManager requires of a fine level).
"""
import numpy as np

import ibamr_amd as ia

# partitions of a 64-cell block edge (drawn uniformly; [64] twice as likely)
_PARTS = [[64], [64], [32, 32], [16, 48], [48, 16], [16, 16, 32], [16, 32, 16], [32, 16, 16], [24, 24, 16],
          [24, 16, 24], [16, 24, 24]]


def shell_points(nshell, per_shell, rng, centre=(0.25, 0.75), axes=(0.05, 0.15)):
    X = []
    for _ in range(nshell):
        c = rng.uniform(centre[0], centre[1], 3)
        ax = rng.uniform(axes[0], axes[1], 3)
        d = rng.standard_normal((per_shell, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        X.append(c + d * ax)
    return np.ascontiguousarray(np.concatenate(X))


def amr_level(coarse=128, ratio=4, npoints=1 << 21, nshell=4, seed=2024, block=64):
    """Returns (level, X): an ibamr_amd.Level of fine patches and float64 points [n][3]."""
    rng = np.random.default_rng(seed)
    N = coarse * ratio
    h = 1.0 / N
    X = shell_points(nshell, npoints // nshell, rng)
    cell = np.floor(X / h).astype(np.int64)
    if cell.min() < 0 or cell.max() >= N:
        raise ValueError("points outside the unit cube")
    nb = N // block
    bid = (cell[:, 0] // block) + nb * ((cell[:, 1] // block) + nb * (cell[:, 2] // block))
    patches = []
    for b in np.unique(bid):
        bz, rem = divmod(int(b), nb * nb)
        by, bx = divmod(rem, nb)
        base = np.array([bx, by, bz]) * block
        sel = cell[bid == b] - base
        segs = [[v * block // 64 for v in _PARTS[rng.integers(0, len(_PARTS))]] for _ in range(3)]
        offs = [np.cumsum([0] + s[:-1]) for s in segs]
        for ox, ex in zip(offs[0], segs[0]):
            for oy, ey in zip(offs[1], segs[1]):
                for oz, ez in zip(offs[2], segs[2]):
                    lo = np.array([ox, oy, oz])
                    hi = lo + np.array([ex, ey, ez])
                    if np.all((sel >= lo) & (sel < hi), axis=1).any():
                        il = (base + lo).tolist()
                        iu = (base + hi - 1).tolist()
                        patches.append(ia.Patch(il, iu, [i * h for i in il], [(i + 1) * h for i in iu]))
    level = ia.Level([h] * 3, [0] * 3, [N - 1] * 3, [0.0] * 3, [1.0] * 3, [0] * 3, patches)
    return level, X
