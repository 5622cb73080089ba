"""GPU Lagrangian forces (ibamr_forces_*) against the oracle's restatement of
IBStandardForceGen::computeLagrangianForce: bitwise in FP64 (every node's force
is the reference loop's sequence of operations), FP32 output within one
rounding of the FP64 value."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import forces_cases as fc  # noqa: E402
from ibamr_amd import structure  # noqa: E402

pytestmark = pytest.mark.gpu


def _big(n=1 << 20, seed=31):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0.2, 0.8, (n, 3))
    i = np.arange(n)
    nx = rng.integers(0, n, n // 2)
    ny = (nx + rng.integers(1, n, n // 2)) % n
    springs = (np.concatenate([i, nx]), np.concatenate([(i + 1) % n, ny]), rng.uniform(1, 10, n + n // 2),
               rng.uniform(0, 1e-3, n + n // 2))
    beams = ((i - 1) % n, i, (i + 1) % n, rng.uniform(0.5, 2.0, n), rng.uniform(-1e-4, 1e-4, (n, 3)))
    tn = np.sort(rng.choice(n, n // 16, replace=False))
    targets = (tn, rng.uniform(1, 10, len(tn)), rng.uniform(0, 1, len(tn)), X[tn] + 1e-3)
    return 3, X, springs, beams, targets


CASES = {"curve2d": fc.curve2d, "ring3d": fc.ring3d, "big": _big}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("accumulate", [False, True])
def test_forces_fp64_bitwise(orc, case, accumulate):
    ndim, X, springs, beams, targets = CASES[case]()
    print("forces", case, "accumulate", accumulate, "n", len(X), flush=True)
    rng = np.random.default_rng(5)
    U = rng.standard_normal(X.shape)
    F0 = rng.standard_normal(X.shape) if accumulate else np.zeros(X.shape)
    want = orc.lagrangian_force(ndim, X, F=F0, U=U, springs=springs, beams=beams, targets=targets)
    f = structure.Forces(ndim, len(X), springs, beams, targets)
    Xd = torch.from_numpy(X).cuda()
    Ud = torch.from_numpy(U).cuda()
    Fd = torch.from_numpy(F0).cuda() if accumulate else torch.full(X.shape, float("nan"), dtype=torch.float64,
                                                                     device="cuda")
    f.compute(Fd, Xd, Ud, accumulate=accumulate)
    got = Fd.cpu().numpy()
    assert np.array_equal(got, want), np.abs(got - want).max()
    # run twice: the same bits
    Fd2 = torch.from_numpy(F0).cuda()
    f.compute(Fd2, Xd, Ud, accumulate=accumulate)
    assert torch.equal(Fd2, torch.from_numpy(got).cuda())
    f.close()


@pytest.mark.parametrize("case", ["ring3d", "big"])
def test_forces_fp32_output(orc, case):
    ndim, X, springs, beams, targets = CASES[case]()
    rng = np.random.default_rng(6)
    F0 = rng.standard_normal(X.shape).astype(np.float32)
    want = orc.lagrangian_force(ndim, X, F=F0.astype(np.float64), springs=springs, beams=beams, targets=targets)
    f = structure.Forces(ndim, len(X), springs, beams, targets)
    Fd = torch.from_numpy(F0).cuda()
    f.compute(Fd, torch.from_numpy(X).cuda(), None, accumulate=True)
    assert np.array_equal(Fd.cpu().numpy(), want.astype(np.float32))
    f.close()


def test_forces_no_velocity_means_no_damping(orc):
    ndim, X, springs, beams, targets = fc.ring3d()
    want = orc.lagrangian_force(ndim, X, targets=targets)
    f = structure.Forces(ndim, len(X), targets=targets)
    Fd = torch.zeros(X.shape, dtype=torch.float64, device="cuda")
    f.compute(Fd, torch.from_numpy(X).cuda())
    assert np.array_equal(Fd.cpu().numpy(), want)
    f.close()


def _big_laws(seed=32):
    """_big with every spring given index 0, 1 or 7 (Hookean, linear strain,
    polynomial with 2 to 5 parameters)."""
    ndim, X, springs, beams, targets = _big()
    rng = np.random.default_rng(seed)
    ns = len(springs[0])
    idx = rng.choice([0, 1, 7], ns).astype(np.int32)
    npar = np.where(idx == 7, rng.integers(2, 6, ns), 2)
    off = np.zeros(ns + 1, dtype=np.int32)
    off[1:] = np.cumsum(npar)
    par = rng.uniform(-20.0, 20.0, int(off[-1]))
    par[off[:-1]] = springs[2]
    par[off[:-1] + 1] = np.where(idx == 1, springs[3] + 1e-3, springs[3])   # rest > 0 for the strain law
    return ndim, X, springs, beams, targets, {0: 0, 1: 1, 7: 2}, (idx, off, par)


@pytest.mark.parametrize("case", ["ring3d_laws", "big_laws"])
def test_spring_laws_fp64_bitwise(orc, case):
    """Springs under registered force functions (IBStandardForceGen::
    registerSpringForceFunction): bitwise against the oracle evaluating the
    same law per spring, with beams and target points beside them."""
    ndim, X, springs, beams, targets, reg, fn = fc.ring3d_laws() if case == "ring3d_laws" else _big_laws()
    print("spring laws", case, "n", len(X), flush=True)
    rng = np.random.default_rng(7)
    U = rng.standard_normal(X.shape)
    want = orc.lagrangian_force(ndim, X, U=U, springs=springs, beams=beams, targets=targets,
                                spring_laws=fc.laws_of(reg, fn))
    f = structure.Forces(ndim, len(X), springs, beams, targets, registry=reg, spring_functions=fn)
    Fd = torch.full(X.shape, float("nan"), dtype=torch.float64, device="cuda")
    f.compute(Fd, torch.from_numpy(X).cuda(), torch.from_numpy(U).cuda())
    got = Fd.cpu().numpy()
    assert np.array_equal(got, want), np.abs(got - want).max()
    f.close()


def test_spring_registry_errors_and_default():
    ndim, X, springs, beams, targets = fc.ring3d()
    import ibamr_amd as ia
    f = structure.Forces(ndim, len(X), springs)
    ns = len(springs[0])
    off = np.arange(0, 2 * ns + 1, 2)
    par = np.stack([springs[2], springs[3]], axis=1).reshape(-1)
    with pytest.raises(ia.IBAMRError, match="no spring force function is registered under index 3"):
        f.set_spring_parameters(np.full(ns, 3), off, par)
    with pytest.raises(ia.IBAMRError, match="unknown spring law"):
        f.register_spring_function(3, 9)
    short = np.arange(ns + 1)
    with pytest.raises(ia.IBAMRError, match="at least stiffness and rest length"):
        f.set_spring_parameters(np.zeros(ns), short, par[:ns])
    # index 0 is the Hookean default: the same parameters through the registry give the created forces
    Xd = torch.from_numpy(X).cuda()
    F1 = torch.zeros(X.shape, dtype=torch.float64, device="cuda")
    f.compute(F1, Xd)
    f.set_spring_parameters(np.zeros(ns), off, par)
    F2 = torch.zeros(X.shape, dtype=torch.float64, device="cuda")
    f.compute(F2, Xd)
    assert torch.equal(F1, F2)
    f.close()
