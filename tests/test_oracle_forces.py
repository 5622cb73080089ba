"""The oracle's restatement of IBStandardForceGen::computeLagrangianForce
(springs, beams, target points) pinned by what the mechanics fixes: forces are
minus the gradient of the stored energy (central differences), internal forces
sum to zero (Newton's third law), coincident spring ends give no force, and the
structure files of the reference's tests read as documented."""
import numpy as np
import pytest

import forces_cases as fc


def _energy(ndim, X, springs, beams, targets):
    E = 0.0
    if springs is not None:
        for m, s, k, r in zip(*springs):
            R = np.linalg.norm(X[s] - X[m])
            E += 0.5 * k * (R - r) ** 2
    if beams is not None:
        for p, m, nx, K, c in zip(*beams):
            d = X[nx] + X[p] - 2 * X[m] - c
            E += 0.5 * K * float(d @ d)
    if targets is not None:
        for t, K, _, x0 in zip(*targets):
            d = x0 - X[t]
            E += 0.5 * K * float(d @ d)
    return E


@pytest.mark.parametrize("case", ["curve2d", "ring3d"])
def test_force_is_minus_energy_gradient(orc, case):
    ndim, X, springs, beams, targets = getattr(fc, case)()
    F = orc.lagrangian_force(ndim, X, springs=springs, beams=beams, targets=targets)
    rng = np.random.default_rng(1)
    eps = 1e-7
    for _ in range(12):
        v = rng.integers(0, X.shape[0])
        d = rng.integers(0, ndim)
        Xp, Xm = X.copy(), X.copy()
        Xp[v, d] += eps
        Xm[v, d] -= eps
        g = (_energy(ndim, Xp, springs, beams, targets) - _energy(ndim, Xm, springs, beams, targets)) / (2 * eps)
        assert abs(F[v, d] + g) <= 1e-5 * max(1.0, abs(g)), (v, d, F[v, d], -g)


@pytest.mark.parametrize("case", ["curve2d", "ring3d"])
def test_internal_forces_sum_to_zero(orc, case):
    ndim, X, springs, beams, targets = getattr(fc, case)()
    F = orc.lagrangian_force(ndim, X, springs=springs, beams=beams)
    scale = np.abs(F).max()
    assert np.abs(F.sum(axis=0)).max() <= 1e-12 * scale * X.shape[0]


def test_target_point_damping_and_accumulation(orc):
    ndim, X, springs, beams, targets = fc.ring3d()
    U = np.random.default_rng(2).standard_normal(X.shape)
    F0 = np.random.default_rng(3).standard_normal(X.shape)
    F = orc.lagrangian_force(ndim, X, F=F0, U=U, targets=targets)
    want = F0.copy()
    tn, K, E, X0 = targets
    want[tn] = F0[tn] + (K[:, None] * (X0 - X[tn]) - E[:, None] * U[tn])
    assert np.array_equal(F, want)


def test_coincident_spring_ends_give_no_force(orc):
    X = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [0.4, 0.2, 0.3]])
    springs = ([0, 1], [1, 2], [5.0, 7.0], [0.5, 0.0])
    F = orc.lagrangian_force(3, X, springs=springs)
    assert np.array_equal(F[0], np.zeros(3))
    assert np.allclose(F[1], [7.0 * 0.3, 0, 0]) and np.allclose(F[2], [-7.0 * 0.3, 0, 0])


def test_curve2d_64_files():
    ndim, X, springs, beams, targets = fc.curve2d()
    assert X.shape == (304, 2) and len(springs[0]) == 304
    assert springs[0][0] == 0 and springs[1][0] == 1
    assert np.all(springs[3] == 0.0) and np.allclose(springs[2], 1.9353241079974475e+02)
