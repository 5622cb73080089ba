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


def _law_energy(law, p, R):
    """Stored energy of one spring whose tension is the law's T(R) (dE/dR = T)."""
    if law == 0:
        return 0.5 * p[0] * (R - p[1]) ** 2
    if law == 1:
        return 0.5 * p[0] * p[1] * (R / p[1] - 1.0) ** 2
    x = R - p[1]
    c = [p[0]] + list(p[2:])
    return sum(cm * x ** (m + 2) / (m + 2) for m, cm in enumerate(c))


def test_spring_laws_force_is_minus_energy_gradient(orc):
    ndim, X, springs, beams, targets, reg, fn = fc.ring3d_laws()
    laws = fc.laws_of(reg, fn)
    assert set(laws[0]) == {0, 1, 2}
    F = orc.lagrangian_force(ndim, X, springs=springs, spring_laws=laws)
    law, off, par = laws

    def energy(Y):
        return sum(_law_energy(law[k], par[off[k]:off[k + 1]], np.linalg.norm(Y[s] - Y[m]))
                   for k, (m, s) in enumerate(zip(springs[0], springs[1])))
    rng = np.random.default_rng(5)
    eps = 1e-7
    for _ in range(12):
        v, d = rng.integers(0, X.shape[0]), rng.integers(0, ndim)
        Xp, Xm = X.copy(), X.copy()
        Xp[v, d] += eps
        Xm[v, d] -= eps
        g = (energy(Xp) - energy(Xm)) / (2 * eps)
        assert abs(F[v, d] + g) <= 1e-5 * max(1.0, abs(g)), (v, d, F[v, d], -g)


def test_spring_laws_closed_forms(orc):
    """Two nodes 0.75 apart along x (exact in binary): linear strain 3 (R/0.5 - 1)
    = 1.5; polynomial {2, 0.25, 3, 4}: x = 0.5, T = 0.5 (2 + 0.5 (3 + 0.5 * 4))
    = 2.25; the master node is pulled towards the slave by T."""
    X = np.array([[0.125, 0.5, 0.5], [0.875, 0.5, 0.5]])
    sp = ([0], [1], [0.0], [0.0])
    for law, p, T in ((1, [3.0, 0.5], 1.5), (2, [2.0, 0.25, 3.0, 4.0], 2.25), (0, [4.0, 0.5], 1.0)):
        F = orc.lagrangian_force(3, X, springs=sp, spring_laws=([law], [0, len(p)], p))
        assert np.array_equal(F, np.array([[T, 0.0, 0.0], [-T, 0.0, 0.0]])), (law, F)


def test_polynomial_with_two_parameters_is_hookean_bitwise(orc):
    ndim, X, springs, beams, targets = fc.ring3d()
    ns = len(springs[0])
    par = np.stack([springs[2], springs[3]], axis=1).reshape(-1)
    off = np.arange(0, 2 * ns + 1, 2)
    want = orc.lagrangian_force(ndim, X, springs=springs, beams=beams, targets=targets)
    got = orc.lagrangian_force(ndim, X, springs=springs, beams=beams, targets=targets,
                               spring_laws=(np.full(ns, 2), off, par))
    assert np.array_equal(got, want)


def test_spring_file_function_index_and_parameters(tmp_path):
    """The .spring columns after the rest length: force function index, then
    further parameters appended to {stiffness, rest} (IBStandardInitializer.cpp:397-434)."""
    from ibamr_amd import structure
    base = str(tmp_path / "s")
    with open(base + ".vertex", "w") as f:
        f.write("3\n0 0\n1 0\n2 0\n")
    with open(base + ".spring", "w") as f:
        f.write("3 # springs\n0 1 5.0 0.5\n1 2 6.0 0.25 7 1.5 -2.0 ! two more\n0 2 1.0 0.0 1\n")
    s = structure.read_structure(base, 2)
    idx, off, par = s["spring_functions"]
    assert list(idx) == [0, 7, 1]
    assert list(off) == [0, 2, 6, 8]
    assert list(par) == [5.0, 0.5, 6.0, 0.25, 1.5, -2.0, 1.0, 0.0]
    assert list(s["springs"][2]) == [5.0, 6.0, 1.0]
