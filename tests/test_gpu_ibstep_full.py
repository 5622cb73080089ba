"""The multi-structure workload at full size: 64 ellipsoidal shells on a
periodic 384^3 side grid, point counts log-uniform in [2000, 200000] (seeded
as the benchmark seeds them, 2.7 M points), shells 0-31 IB_4 and 32-63
BSPLINE_3, each kernel group one level plan (IBStrategySet: the first group
overwrites f, the second accumulates; each interpolates its own points).

  * one eager FP64 step against the oracle, bit for bit: f after both
    groups' spreads (the oracle run over each group's entries in the library's
    bin order), U interpolated from the step's own u (the stand-in fluid
    update u = u_bg + alpha f is a torch kernel, not the library's), and
    X + dt U with the periodic wrap;
  * captured-graph replays equal the eager steps bitwise;
  * after 100 graph-replayed steps (dt a quarter cell at the largest velocity)
    the FP32C positions stay within 2e-4 cells of the FP64 ones."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from ibamr_amd import ibstep  # noqa: E402

pytestmark = pytest.mark.gpu
N, NSHELL = 384, 64


def _workload():
    rng = np.random.default_rng(55)
    counts = [int(round(math.exp(v))) for v in rng.uniform(math.log(2000), math.log(200000), NSHELL)]
    structs = [("IB_4" if i < NSHELL // 2 else "BSPLINE_3", c) for i, c in enumerate(counts)]
    X0 = ibstep.ellipsoid_shells(NSHELL, counts, seed=56)
    g = torch.Generator(device="cuda")
    g.manual_seed(57)
    F64 = torch.randn(X0.shape, generator=g, device="cuda", dtype=torch.float64)
    return structs, X0, F64


def _set(level, structs, precision):
    h = 1.0 / N
    dt_ = torch.float64 if precision == ia.FP64 else torch.float32
    gw = max(ia.get_minimum_ghost_width(k) for k, _ in structs)
    return ibstep.StructureSet(level, structs, precision, alpha=0.05 * h ** 3,
                               u_bg=ibstep.background_velocity(level, [gw] * 3, dt_)), dt_


def test_full_size_step_oracle_graph_and_drift(orc):
    level = ia.Level.uniform_periodic(N, 3)
    h = 1.0 / N
    structs, X0, F64 = _workload()
    n = X0.shape[0]
    print("multi-structure full size:", n, "points", flush=True)
    S, _ = _set(level, structs, ia.FP64)
    X = X0.clone()
    U = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    # dt as the benchmark sets it: a quarter cell at the largest velocity of a first step
    S.step(X, F64, U, 0.0)
    torch.cuda.synchronize()
    dt = 0.25 * h / float(U.abs().max())
    S.step(X, F64, U, dt)
    torch.cuda.synchronize()
    # ---- the step against the oracle
    gw = S.gcw[0]
    il, iu = [0] * 3, [N - 1] * 3
    Xn, Fn = X0.cpu().numpy(), F64.cpu().numpy()
    ref = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    bounds = [(a, a + m) for _, a, m in S.groups]
    for (k, _, _), (a, b) in zip(S.groups, bounds):
        Xs, Fs = np.ascontiguousarray(Xn[a:b]), np.ascontiguousarray(Fn[a:b])
        cells = orc.get_cell_index(Xs, [0.0] * 3, [1.0] * 3, [h] * 3, il, iu)
        orc.spread_patch(k, True, [h] * 3, [0.0] * 3, il, iu, [gw] * 3, ref, Xs, Fs, cells,
                         ia.get_minimum_ghost_width(k), [N] * 3, il)
    for d in range(3):
        assert np.array_equal(S.f[d].cpu().numpy(), ref[d]), d
    del ref
    u = [np.ascontiguousarray(S.u[d].cpu().numpy()) for d in range(3)]
    Ug = U.cpu().numpy()
    for (k, _, _), (a, b) in zip(S.groups, bounds):
        Uo = orc.interp_side(k, [h] * 3, [0.0] * 3, il, iu, [gw] * 3, u, np.ascontiguousarray(Xn[a:b]))
        assert np.array_equal(Ug[a:b], Uo), k
    Xw = Xn + dt * Ug
    Xw = np.where(Xw < 0.0, Xw + 1.0, Xw)
    Xw = np.where(Xw >= 1.0, Xw - 1.0, Xw)
    assert np.array_equal(X.cpu().numpy(), Xw)
    del u
    # ---- graph replays against eager steps (bitwise)
    traj = [X.clone()]
    for _ in range(2):
        S.step(X, F64, U, dt)
        traj.append(X.clone())
    torch.cuda.synchronize()
    X.copy_(X0)
    torch.cuda.synchronize()
    g = S.capture(X, F64, U, dt)
    torch.cuda.synchronize()
    X.copy_(X0)
    torch.cuda.synchronize()
    for k in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(X, traj[k]), k
    del g
    # ---- FP32C drift against FP64 over 100 graph-replayed steps
    out = {}
    for prec in (ia.FP64, ia.FP32C):
        Sp, dt_ = _set(level, structs, prec)
        Xp = X0.clone()
        Fp = F64.to(dt_).contiguous()
        Up = torch.zeros((n, 3), dtype=dt_, device="cuda")
        Sp.step(Xp, Fp, Up, dt)
        torch.cuda.synchronize()
        Xp.copy_(X0)
        torch.cuda.synchronize()
        gp = Sp.capture(Xp, Fp, Up, dt)
        torch.cuda.synchronize()
        Xp.copy_(X0)
        torch.cuda.synchronize()
        for _ in range(100):
            gp.replay()
        torch.cuda.synchronize()
        out[prec] = Xp
        del gp, Sp
    d = (out[ia.FP32C] - out[ia.FP64]).abs()
    drift = float(torch.minimum(d, 1.0 - d).max()) / h   # periodic distance
    print("FP32C drift after 100 steps: %.3g cells" % drift, flush=True)
    assert drift < 2e-4
