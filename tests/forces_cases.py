"""Structures for the force tests: the reference's tests/IB/curve2d_64 input
files (fixtures) and a seeded 3-D structure with springs, beams and target
points."""
import os

import numpy as np

from ibamr_amd import structure

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def curve2d():
    s = structure.read_structure(os.path.join(GOLD, "curve2d_64"), 2)
    rng = np.random.default_rng(4)
    X = s["X"] + rng.uniform(-1e-3, 1e-3, s["X"].shape)   # stretched a little
    return 2, X, s["springs"], None, None


def ring3d(n=500, seed=9):
    rng = np.random.default_rng(seed)
    t = 2 * np.pi * np.arange(n) / n
    X0 = np.stack([0.5 + 0.2 * np.cos(t), 0.5 + 0.2 * np.sin(t), 0.5 + 0.01 * np.sin(3 * t)], axis=1)
    X = X0 + rng.uniform(-2e-3, 2e-3, X0.shape)
    i = np.arange(n)
    springs = (i, (i + 1) % n, rng.uniform(50, 150, n), np.full(n, 2 * np.pi * 0.2 / n))
    # a few extra springs across the ring, one of them with coincident ends after a move
    extra_m = rng.integers(0, n, 40)
    extra_s = (extra_m + rng.integers(5, n // 2, 40)) % n
    springs = (np.concatenate([springs[0], extra_m]), np.concatenate([springs[1], extra_s]),
               np.concatenate([springs[2], rng.uniform(1, 5, 40)]), np.concatenate([springs[3], np.zeros(40)]))
    beams = ((i - 1) % n, i, (i + 1) % n, rng.uniform(0.5, 2.0, n), rng.uniform(-1e-4, 1e-4, (n, 3)))
    tn = np.arange(0, n, 7)
    targets = (tn, rng.uniform(100, 1000, len(tn)), rng.uniform(0.0, 2.0, len(tn)), X0[tn])
    return 3, X, springs, beams, targets


def ring3d_laws(seed=10):
    """ring3d with the springs spread over three force function indices:
    0 Hookean {kappa, rest}, 1 linear strain {stiffness, rest > 0}, 7 a
    polynomial in the extension with 2 to 5 parameters."""
    ndim, X, springs, beams, targets = ring3d()
    rng = np.random.default_rng(seed)
    ns = len(springs[0])
    idx = rng.choice([0, 1, 7], ns).astype(np.int32)
    par = []
    for k in range(ns):
        rest = float(springs[3][k])
        if idx[k] == 0:
            par.append([float(springs[2][k]), rest])
        elif idx[k] == 1:
            par.append([float(springs[2][k]), rest if rest > 0 else 0.01])
        else:
            par.append([float(springs[2][k]), rest] + list(rng.uniform(-50.0, 50.0, rng.integers(0, 4))))
    off = np.zeros(ns + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(p) for p in par])
    registry = {0: 0, 1: 1, 7: 2}
    return ndim, X, springs, beams, targets, registry, (idx, off, np.array([v for p in par for v in p]))


def laws_of(registry, functions):
    """(law per spring, offsets, params) from a registry and the per-spring indices."""
    idx, off, par = functions
    return np.array([registry[int(i)] for i in idx], dtype=np.int32), off, par
