"""The allocation-free interpolation (ibamr_s0_interp_noalloc, device
pointers, scratch in a file-scope device array) against the allocating form
ibamr_s0_interp on the same inputs: bitwise equal on two calls in a row, the
second with fewer points (the scratch keeps the first call's contents and must
be cleared), with repeated indices and periodic-image shifts; a point count
above the scratch capacity is refused before any GPU work."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import ibamr_amd as ia  # noqa: E402

pytestmark = pytest.mark.gpu
IBAMR_ERR_CAPACITY = -6
MAX_POINTS = 1 << 22

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)


def _fn():
    f = ia.lib().ibamr_s0_interp_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p, ctypes.c_int, _d, _d, ctypes.c_int, _i, _i, _i, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_void_p]
    return f


def _arr(v, t):
    a = np.ascontiguousarray(v, dtype=np.float64 if t is ctypes.c_double else np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(t))


def _case(ndim, n, kernel, npts, depth, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    gw = ia.get_minimum_ghost_width(kernel)
    il, iu = [2] * ndim, [2 + n - 1] * ndim
    xl = [2 * h] * ndim
    X = rng.uniform(xl[0] - (gw + 1.5) * h, xl[0] + (n + gw + 1.5) * h, (npts, ndim))
    idx = rng.permutation(npts).astype(np.int32)
    idx = np.concatenate([idx, idx[: npts // 5]])              # repeated indices: the last entry wins
    sh = (rng.integers(-1, 2, (len(idx), ndim)) * (n * h)).astype(np.float64)
    shp = tuple([depth] + [iu[d] - il[d] + 1 + 2 * gw for d in reversed(range(ndim))])
    u = rng.uniform(-1, 1, shp)
    V0 = rng.uniform(-1, 1, (npts, depth))
    return h, gw, il, iu, xl, X, idx, sh, u, V0


def _both(kernel, ndim, n, npts, depth, seed):
    print("no-alloc case:", kernel, ndim, n, npts, depth, flush=True)
    h, gw, il, iu, xl, X, idx, sh, u, V0 = _case(ndim, n, kernel, npts, depth, seed)
    Vref = V0.copy()
    ia.s0_interp(kernel, [h] * ndim, xl, il, iu, [gw] * ndim, u, X, Vref, idx, sh, depth=depth)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to("cuda") for k, v in
           dict(u=u, idx=idx, sh=sh, X=X, V=V0).items()}
    keep = [_arr([h] * ndim, ctypes.c_double), _arr(xl, ctypes.c_double), _arr(il, ctypes.c_int),
            _arr(iu, ctypes.c_int), _arr([gw] * ndim, ctypes.c_int)]
    rc = _fn()(kernel.encode(), ndim, keep[0][1], keep[1][1], depth, keep[2][1], keep[3][1], keep[4][1],
               dev["u"].data_ptr(), dev["idx"].data_ptr(), dev["sh"].data_ptr(), len(idx), dev["X"].data_ptr(),
               npts, dev["V"].data_ptr(), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, ia.lib().ibamr_last_error()
    torch.cuda.synchronize()
    got = dev["V"].cpu().numpy()
    assert np.array_equal(got, Vref)
    assert not np.array_equal(Vref, V0)


def test_noalloc_matches_allocating_form_twice():
    _both("IB_6", 3, 12, 3000, 2, 31)      # first call
    _both("IB_6", 3, 12, 700, 2, 32)       # second call, fewer points: the scratch holds the first call's entries
    _both("BSPLINE_3", 2, 24, 400, 1, 33)  # another kernel and dimension


def test_noalloc_capacity_refused():
    keep = [_arr([0.1] * 3, ctypes.c_double), _arr([0.0] * 3, ctypes.c_double), _arr([0] * 3, ctypes.c_int),
            _arr([3] * 3, ctypes.c_int), _arr([3] * 3, ctypes.c_int)]
    rc = _fn()(b"IB_4", 3, keep[0][1], keep[1][1], 1, keep[2][1], keep[3][1], keep[4][1], None, None, None, 10,
               None, MAX_POINTS + 1, None, None)
    assert rc == IBAMR_ERR_CAPACITY
    assert b"scratch" in ia.lib().ibamr_last_error()
