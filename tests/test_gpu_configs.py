"""GPU parity at the full sizes of the two FP64 workloads:

  C1: 2-D periodic unit square, one 128^2 patch, side-centred, IB_4; 1,024
      points X_k = (0.5 + 0.25 cos t_k, 0.5 + 0.25 sin t_k), t_k = 2 pi k / 1024;
      F ~ N(0, 1); face values u ~ U[-1, 1].
  C2: 3-D periodic 256^3, side-centred, IB_4, 2^20 points on a Fibonacci sphere
      of radius 0.3 about the centre, each coordinate jittered by U[-h/4, h/4];
      F ~ N(0, 1); u ~ U[-1, 1].

Both FP64: spread (accumulating into a random field) and interp bitwise equal
to the oracle run over the same entry lists in the library's summation order.
Seeds: numpy default_rng(101) for C1, default_rng(202) for C2.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _run(orc, ndim, N, kernel, X, F, seed):
    rng = np.random.default_rng(seed)
    level = ia.Level.uniform_periodic(N, ndim)
    h = 1.0 / N
    plan = ia.LevelPlan(level, kernel, ia.SIDE, ia.FP64, max_points=X.shape[0])
    gw = plan.data_gcw
    f0 = [rng.uniform(-1, 1, ia.array_shape(level, 0, gw, ia.SIDE, axis=a)) for a in range(ndim)]
    u = [rng.uniform(-1, 1, ia.array_shape(level, 0, gw, ia.SIDE, axis=a)) for a in range(ndim)]
    Xd, Fd = _dev(X), _dev(F)
    print("config", ndim, N, kernel, X.shape[0], flush=True)
    plan.bin(Xd)
    fd = [_dev(a) for a in f0]
    plan.bind(0, fd)
    plan.spread(0, Fd, Xd, mode=ia.ACCUMULATE)
    ud = [_dev(a) for a in u]
    plan.bind(1, ud)
    plan.fill_periodic_ghosts(1)
    U = torch.zeros((X.shape[0], ndim), dtype=torch.float64, device="cuda")
    plan.interp(1, U, Xd)
    torch.cuda.synchronize()
    il, iu = [0] * ndim, [N - 1] * ndim
    cells = orc.get_cell_index(X, [0.0] * ndim, [1.0] * ndim, [h] * ndim, il, iu)
    ref = [a.copy() for a in f0]
    orc.spread_patch(kernel, True, [h] * ndim, [0.0] * ndim, il, iu, gw, ref, X, F, cells, gw[0], [N] * ndim, il)
    for a in range(ndim):
        assert np.array_equal(fd[a].cpu().numpy(), ref[a]), ("spread axis", a)
    uf = [t.cpu().numpy() for t in ud]
    tile = 8 if ndim == 3 else 16
    idx, sh = orc.build_patch_entries(cells, il, iu, 0, [N] * ndim, il, [h] * ndim, [tile] * ndim,
                                      interior_only=True)
    assert len(idx) == X.shape[0]
    Uref = orc.interp_side(kernel, [h] * ndim, [0.0] * ndim, il, iu, gw, uf, X, indices=idx, Xshift=sh)
    assert np.array_equal(U.cpu().numpy(), Uref)


def test_config1_2d_circle(orc):
    n = 1024
    t = 2 * math.pi * np.arange(n) / n
    X = np.stack([0.5 + 0.25 * np.cos(t), 0.5 + 0.25 * np.sin(t)], axis=1)
    F = np.random.default_rng(101).standard_normal((n, 2))
    _run(orc, 2, 128, "IB_4", X, F, 101)


def test_config2_3d_fibonacci_sphere(orc):
    n, N = 1 << 20, 256
    h = 1.0 / N
    rng = np.random.default_rng(202)
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    r = np.sqrt(1 - z * z)
    phi = math.pi * (3 - math.sqrt(5)) * np.arange(n)
    X = 0.5 + 0.3 * np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    X += rng.uniform(-h / 4, h / 4, X.shape)
    F = rng.standard_normal((n, 3))
    _run(orc, 3, N, "IB_4", X, F, 203)
