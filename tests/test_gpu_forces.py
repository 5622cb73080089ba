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
