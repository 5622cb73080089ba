"""GPU tests of the multi-structure timestep (ibamr_amd.ibstep).

  * one eager step: f = S_IB4 F_a + S_BSPLINE3 F_b (two plans spreading into
    one field, the second accumulating), against the oracle to 1e-13;
  * a captured HIP graph replayed several times, with eager work and
    synchronizations between replays, reproduces the eager trajectory bitwise
    (FP64 kernels are deterministic);
  * the FP32C trajectory stays within 1e-4 h of the FP64 one after 10 steps.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402
from ibamr_amd import ibstep  # noqa: E402

pytestmark = pytest.mark.gpu
N = 48
STRUCT = [("IB_4", 1500), ("IB_4", 700), ("BSPLINE_3", 2000), ("BSPLINE_3", 900)]


def _setup(precision, seed=5):
    level = ia.Level.uniform_periodic(N, 3)
    counts = [n for _, n in STRUCT]
    X = ibstep.ellipsoid_shells(len(STRUCT), counts, seed=seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed + 1)
    F64 = torch.randn(X.shape, generator=g, device="cuda", dtype=torch.float64)
    h = 1.0 / N
    gw = max(ia.get_minimum_ghost_width(k) for k, _ in STRUCT)
    dt_ = torch.float64 if precision == ia.FP64 else torch.float32
    ubg = ibstep.background_velocity(level, [gw] * 3, dt_)
    S = ibstep.StructureSet(level, STRUCT, precision, alpha=0.05 * h ** 3, u_bg=ubg)
    F = F64.to(dt_).contiguous()
    U = torch.zeros(X.shape, dtype=dt_, device="cuda")
    return S, X, F, U, F64


def test_two_kernel_spread_matches_oracle(orc):
    S, X, F, U, _ = _setup(ia.FP64)
    X0 = X.clone()
    print("ibstep eager step", flush=True)
    S.step(X, F, U, dt=0.0)
    torch.cuda.synchronize()
    h = 1.0 / N
    gw = S.gcw[0]
    il, iu = [0] * 3, [N - 1] * 3
    Xn, Fn = X0.cpu().numpy(), F.cpu().numpy()
    ref = [np.zeros(orc.side_axis_shape(il, iu, [gw] * 3, a)) for a in range(3)]
    start = 0
    for k, n in STRUCT:
        Xs, Fs = Xn[start:start + n], Fn[start:start + n]
        cells = orc.get_cell_index(Xs, [0.0] * 3, [1.0] * 3, [h] * 3, il, iu)
        idx, sh = orc.build_patch_entries(cells, il, iu, gw, [N] * 3, il, [h] * 3, [8] * 3)
        orc.spread_side(k, [h] * 3, [0.0] * 3, il, iu, [gw] * 3, ref, Xs, Fs, indices=idx, Xshift=sh)
        start += n
    for a in range(3):
        got = S.f[a].cpu().numpy()
        sl = tuple(slice(gw, gw + N + (1 if (2 - ax) == a else 0)) for ax in range(3))
        err = np.max(np.abs(got[sl] - ref[a][sl]))
        assert err <= 1e-13 * np.max(np.abs(ref[a][sl])), (a, err)
    # dt = 0: positions unchanged, U = J(u_bg + alpha f) is finite
    assert torch.equal(X, X0)
    assert torch.isfinite(U).all()


def test_graph_replays_match_eager_bitwise():
    S, X, F, U, _ = _setup(ia.FP64)
    X0 = X.clone()
    dt = 0.25 / N / 0.1
    steps = 4
    traj = []
    for _ in range(steps):
        S.step(X, F, U, dt)
        traj.append(X.clone())
    torch.cuda.synchronize()
    assert not torch.equal(traj[-1], X0)
    # warm-up done above; capture one step, then replay from X0
    X.copy_(X0)
    torch.cuda.synchronize()
    print("ibstep capture", flush=True)
    g = S.capture(X, F, U, dt)
    torch.cuda.synchronize()
    X.copy_(X0)
    torch.cuda.synchronize()
    for k in range(steps):
        g.replay()
        torch.cuda.synchronize()
        # eager work between replays
        junk = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        assert torch.equal(X, traj[k]), k


def test_autotune_then_graph_matches_untuned_fp64():
    """StructureSet.autotune times every built tile shape per kernel group and
    keeps the fastest; FP64 spreads are bitwise the same with every shape, so a
    captured step after it gives the trajectory of an untuned set, bit for bit.
    One group is moved off its default shape before capture, so the captured
    step runs a non-default work list."""
    S0, X, F, U0, _ = _setup(ia.FP64)
    S1, _, _, U1, _ = _setup(ia.FP64)
    X0 = X.clone()
    dt = 0.25 / N / 0.1
    X1 = X.clone()
    S1.step(X1, F, U1, dt)  # warm-up (bins) before autotune and capture
    torch.cuda.synchronize()
    X1.copy_(X0)
    tunes = S1.autotune(X1, F, reps=1)
    torch.cuda.synchronize()
    assert sorted(tunes) == ["BSPLINE_3", "IB_4"]
    for (k, _, _), p in zip(S1.groups, S1.plans):
        assert sorted(tunes[k]) == [(8, 8, 16), (8, 8, 32), (16, 4, 32)]
        assert p.tile_shape == min(tunes[k], key=tunes[k].get)
    p0 = S1.plans[0]
    p0.tile_shape = (16, 4, 32) if p0.tile_shape != (16, 4, 32) else (8, 8, 32)
    g = S1.capture(X1, F, U1, dt)
    torch.cuda.synchronize()
    X1.copy_(X0)
    torch.cuda.synchronize()
    for k in range(3):
        S0.step(X, F, U0, dt)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(U1, U0), k
        assert torch.equal(X1, X), k


def test_fp32c_trajectory_tracks_fp64():
    S64, X64, F64t, U64, _ = _setup(ia.FP64)
    S32, X32, F32t, U32, _ = _setup(ia.FP32C)
    assert torch.equal(X64, X32)
    dt = 0.25 / N / 0.1
    for _ in range(10):
        S64.step(X64, F64t, U64, dt)
        S32.step(X32, F32t, U32, dt)
    torch.cuda.synchronize()
    drift = float((X64 - X32).abs().max()) * N
    assert drift < 1e-4, drift


def _ring_forces(npoints, seed=12):
    """Springs joining consecutive points of each structure into a closed loop,
    beams along them and target points at every 5th point."""
    from ibamr_amd import structure
    rng = np.random.default_rng(seed)
    m, s, pv, nx = [], [], [], []
    start = 0
    for _, n in STRUCT:
        i = np.arange(n)
        m.append(start + i)
        s.append(start + (i + 1) % n)
        pv.append(start + (i - 1) % n)
        nx.append(start + (i + 1) % n)
        start += n
    m, s, pv, nx = map(np.concatenate, (m, s, pv, nx))
    springs = (m, s, rng.uniform(0.5, 2.0, len(m)), np.full(len(m), 1e-3))
    beams = (pv, m, nx, rng.uniform(1e-3, 1e-2, len(m)), np.zeros((len(m), 3)))
    tn = np.arange(0, npoints, 5)
    return springs, beams, tn, rng, structure


def test_step_with_force_generation_graph_and_oracle(orc):
    """F generated on the GPU inside the step: the first eager step's F equals
    the oracle's forces bitwise, and a captured graph of the whole step
    (force -> spread -> ... -> update) replays the eager trajectory bitwise."""
    S, X, F, U, _ = _setup(ia.FP64)
    springs, beams, tn, rng, structure = _ring_forces(X.shape[0])
    Xn0 = X.cpu().numpy()
    targets = (tn, rng.uniform(1.0, 5.0, len(tn)), rng.uniform(0.0, 1.0, len(tn)), Xn0[tn])
    S.forces = structure.Forces(3, X.shape[0], springs, beams, targets)
    U.copy_(torch.from_numpy(rng.standard_normal(U.shape) * 1e-2))
    U0 = U.cpu().numpy()
    X0 = X.clone()
    dt = 0.25 / N / 0.1
    print("ibstep with forces: eager", flush=True)
    S.step(X, F, U, dt)
    torch.cuda.synchronize()
    want = orc.lagrangian_force(3, Xn0, U=U0, springs=springs, beams=beams, targets=targets)
    assert np.array_equal(F.cpu().numpy(), want)
    traj = [X.clone()]
    for _ in range(3):
        S.step(X, F, U, dt)
        traj.append(X.clone())
    torch.cuda.synchronize()
    X.copy_(X0)
    U.copy_(torch.from_numpy(U0))
    torch.cuda.synchronize()
    print("ibstep with forces: capture", flush=True)
    g = S.capture(X, F, U, dt)
    torch.cuda.synchronize()
    X.copy_(X0)
    U.copy_(torch.from_numpy(U0))
    torch.cuda.synchronize()
    for k in range(4):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(X, traj[k]), k
    S.forces.close()


def test_points_crossing_the_periodic_boundary_stay_in_the_spread():
    """Points that move across x = 1 are wrapped back into the box inside the
    step (as the reference does at a regrid), so they stay binned: the next
    step's spread still holds every point's force (sum of f h^3 over the
    periodic faces equals sum F, partition of unity)."""
    level = ia.Level.uniform_periodic(N, 3)
    h = 1.0 / N
    n = 500
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    X = torch.rand((n, 3), generator=g, device="cuda", dtype=torch.float64)
    X[:, 0] = 1.0 - 0.3 * h * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)
    gw = ia.get_minimum_ghost_width("IB_4")
    ubg = [torch.full(ia.array_shape(level, 0, [gw] * 3, ia.SIDE, axis=a), 0.5 if a == 0 else 0.0,
                      dtype=torch.float64, device="cuda") for a in range(3)]
    S = ibstep.StructureSet(level, [("IB_4", n)], ia.FP64, alpha=0.0, u_bg=ubg)
    F = torch.randn((n, 3), generator=g, device="cuda", dtype=torch.float64)
    U = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    print("ibstep periodic crossing", flush=True)
    S.step(X, F, U, dt=h)
    torch.cuda.synchronize()
    assert torch.allclose(U[:, 0], torch.full_like(U[:, 0], 0.5), rtol=0, atol=1e-12)
    assert bool(((X >= 0.0) & (X < 1.0)).all())
    assert int((X[:, 0] < 0.5).sum()) > n // 2  # most points crossed and were wrapped
    S.step(X, F, U, dt=0.0)
    torch.cuda.synchronize()
    for a in range(3):
        sl = tuple(slice(gw, gw + N) for _ in range(3))
        tot = float(S.f[a][sl].sum()) * h ** 3
        ref = float(F[:, a].sum())
        assert abs(tot - ref) <= 1e-10 * max(1.0, abs(ref)), (a, tot, ref)
