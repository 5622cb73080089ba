"""The benchmark workload at its full size (3-D periodic 512^3, side-centred,
IB_6, 2^23 points: half uniform, half in a shell of radius 0.3 and thickness
2h), checked through size-independent properties, since the oracle does not
finish at this size in seconds:

  * adjoint identity h^3 <S F, u> = <F, J u> (FP64, spread and interp of the
    same binning; u periodic),
  * the kernel's partition of unity: h^3 sum of every spread component = sum F,
  * linearity: S(2F) = 2 S(F) bit for bit (scaling by 2 is exact),
  * the FP32-compensated form within 1e-6 relative of FP64 (spread and interp).
"""
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu
N, NPTS, KERNEL = 512, 1 << 23, "IB_6"


def _points():
    g = torch.Generator(device="cuda")
    g.manual_seed(2026)
    h = 1.0 / N
    nu = NPTS // 2
    Xu = torch.rand((nu, 3), generator=g, device="cuda", dtype=torch.float64)
    d = torch.randn((NPTS - nu, 3), generator=g, device="cuda", dtype=torch.float64)
    d = d / d.norm(dim=1, keepdim=True)
    r = 0.3 - h + 2 * h * torch.rand((NPTS - nu, 1), generator=g, device="cuda", dtype=torch.float64)
    X = torch.remainder(torch.cat([Xu, 0.5 + r * d]), 1.0).contiguous()
    F = torch.randn((NPTS, 3), generator=g, device="cuda", dtype=torch.float64)
    return X, F


def test_full_size_properties():
    print("full size: points", flush=True)
    X, F = _points()
    level = ia.Level.uniform_periodic(N, 3)
    h = 1.0 / N
    res = {}
    for prec, dt in ((ia.FP64, torch.float64), (ia.FP32C, torch.float32)):
        print("full size: precision", prec, flush=True)
        plan = ia.LevelPlan(level, KERNEL, ia.SIDE, prec, max_points=NPTS)
        gw = plan.data_gcw
        shp = [ia.array_shape(level, 0, gw, ia.SIDE, axis=a) for a in range(3)]
        f = [torch.empty(s, dtype=dt, device="cuda") for s in shp]
        gen = torch.Generator(device="cuda")
        gen.manual_seed(77)
        u = [(2 * torch.rand(s, generator=gen, device="cuda", dtype=torch.float64) - 1).to(dt) for s in shp]
        for a in range(3):
            # face N of axis a is the periodic image of face 0: make u single-valued
            dim = 2 - a
            u[a].narrow(dim, gw[a] + N, 1).copy_(u[a].narrow(dim, gw[a], 1))
        plan.bind(0, f)
        plan.bind(1, u)
        plan.fill_periodic_ghosts(1)
        plan.bin(X)
        Fd = F.to(dt).contiguous()
        plan.spread(0, Fd, X, mode=ia.OVERWRITE)
        U = torch.empty((NPTS, 3), dtype=dt, device="cuda")
        plan.interp(1, U, X)
        torch.cuda.synchronize()
        res[prec] = ([t.double() for t in f], U.double(), [t.double() for t in u])
        if prec == ia.FP64:
            f2 = [torch.empty_like(t) for t in f]
            plan.bind(2, f2)
            plan.spread(2, (2.0 * Fd).contiguous(), X, mode=ia.OVERWRITE)
            torch.cuda.synchronize()
            for a in range(3):
                assert torch.equal(f2[a], 2.0 * f[a]), ("linearity", a)
            del f2
        del plan, f, u, U
        torch.cuda.empty_cache()
    f64, U64, u64 = res[ia.FP64]
    gw = (f64[1].shape[2] - N) // 2
    lhs = 0.0
    for a in range(3):
        sl = (slice(gw, gw + N),) * 3
        fa, ua = f64[a][sl], u64[a][sl]
        lhs += float((fa * ua).sum()) * h ** 3
        s = float(fa.sum()) * h ** 3
        assert abs(s - float(F[:, a].sum())) <= 1e-9 * float(F[:, a].abs().sum()), ("partition of unity", a)
    rhs = float((F * U64).sum())
    assert abs(lhs - rhs) <= 1e-10 * float((F.abs() * U64.abs()).sum()), ("adjoint", lhs, rhs)
    f32, U32, _ = res[ia.FP32C]
    for a in range(3):
        err = float((f32[a] - f64[a]).abs().max() / f64[a].abs().max())
        assert err <= 1e-6, ("fp32c spread", a, err)
    err = float((U32 - U64).abs().max() / U64.abs().max())
    assert err <= 1e-6, ("fp32c interp", err)
