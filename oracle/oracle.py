"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module; the product package never does.  See le_oracle.h for what each
function restates (reference file:line) and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KERNELS = {"PIECEWISE_LINEAR": 0, "IB_4": 1, "IB_6": 2, "BSPLINE_3": 3, "BSPLINE_4": 4, "BSPLINE_5": 5, "BSPLINE_6": 6}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        # IBAMR_ORACLE_LIB: another build of the same sources (the bench's
        # host-tuned CPU baseline, `make native`)
        path = os.environ.get("IBAMR_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_kernel_from_name.argtypes = [ctypes.c_char_p]
        L.orc_stencil_size.argtypes = [ctypes.c_int]
        L.orc_min_ghost_width.argtypes = [ctypes.c_int]
        L.orc_kernel_width.argtypes = [ctypes.c_int]
        L.orc_ib_4_delta.argtypes = [ctypes.c_double]
        L.orc_ib_4_delta.restype = ctypes.c_double
        L.orc_bspline_3_delta.argtypes = [ctypes.c_double]
        L.orc_bspline_3_delta.restype = ctypes.c_double
        L.orc_weights_1d.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_weights_1d_u.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp]
        L.orc_delta.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_delta.restype = ctypes.c_double
        L.orc_interp.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _dp,
                                 _ip, _dp, ctypes.c_int, _dp, _dp]
        L.orc_spread.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                 _dp, _dp, _ip, _ip, _ip, _dp]
        L.orc_get_cell_index.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _ip, _ip]
        L.orc_get_cell_indices.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _ip, _ip, _ip]
        L.orc_build_local_indices.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _ip, _ip, _dp,
                                              ctypes.c_int, _ip]
        L.orc_interp_side.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _ip, _ip,
                                      ctypes.POINTER(_dp), _ip, _dp, ctypes.c_int, _dp, _dp]
        L.orc_spread_side.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _ip, _ip,
                                      ctypes.POINTER(_dp), _ip, _dp, ctypes.c_int, _dp, _dp]
        L.orc_build_patch_entries.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip,
                                              _ip, _dp, _ip, ctypes.c_int, ctypes.c_int, _ip, _dp]
        L.orc_order_entries_z.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, _dp, _dp, _ip, _ip,
                                          ctypes.c_int, _lp, _ip]
        L.orc_bin_ranks.argtypes = [_ip, ctypes.c_int, _ip, _ip, _lp]
        L.orc_bruteforce_spread_at.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                               ctypes.c_int]
        L.orc_bruteforce_spread_at.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _dv(x, n=None):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return a


def _iv(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def kernel_id(name):
    return KERNELS[name] if isinstance(name, str) else int(name)


def stencil_size(kernel):
    return lib().orc_stencil_size(kernel_id(kernel))


def min_ghost_width(kernel):
    return lib().orc_min_ghost_width(kernel_id(kernel))


def kernel_width(kernel):
    return lib().orc_kernel_width(kernel_id(kernel))


def ib_4_delta(r):
    return lib().orc_ib_4_delta(float(r))


def bspline_3_delta(x):
    return lib().orc_bspline_3_delta(float(x))


def delta(kernel, r):
    """The kernel's scalar delta function (IB_4 and the B-splines)."""
    return lib().orc_delta(kernel_id(kernel), float(r))


def weights_1d(kernel, Xs, x_lower, dx, Xu=None):
    """Returns (first stencil index relative to ilower, weights).  Xu: the
    unshifted X, which BSPLINE_4/6 compare with the cell centre (default Xs)."""
    k = kernel_id(kernel)
    w = np.zeros(8)
    lo = lib().orc_weights_1d_u(k, float(Xs), float(Xs if Xu is None else Xu), float(x_lower), float(dx), _d(w))
    return lo, w[: kernel_width(k)].copy()


def ghosted_shape(ilower, iupper, gcw, depth=1):
    """numpy shape (depth, n2, n1, n0) of a column-major ghosted patch array."""
    n = [iupper[d] - ilower[d] + 1 + 2 * gcw[d] for d in range(len(ilower))]
    return (depth,) + tuple(reversed(n))


def interp(kernel, dx, x_lower, ilower, iupper, nugc, u, X, indices=None, Xshift=None, depth=1):
    """Fortran-equivalent interp: u is (depth, [n2,] n1, n0); X is (N, ndim). Returns V (N, depth)
    with untouched rows left as NaN."""
    ndim = len(ilower)
    X = _dv(X).reshape(-1, ndim)
    N = X.shape[0]
    if indices is None:
        indices = np.arange(N)
    indices = _iv(indices)
    if Xshift is None:
        Xshift = np.zeros((len(indices), ndim))
    Xshift = _dv(Xshift)
    V = np.full((N, depth), np.nan)
    u = _dv(u)
    lib().orc_interp(kernel_id(kernel), ndim, _d(_dv(dx)), _d(_dv(x_lower)), depth, _i(_iv(ilower)),
                     _i(_iv(iupper)), _i(_iv(nugc)), _d(u), _i(indices), _d(Xshift), len(indices), _d(X), _d(V))
    return V


def spread(kernel, dx, x_lower, ilower, iupper, nugc, u, X, V, indices=None, Xshift=None, depth=1):
    """Fortran-equivalent spread: accumulates into u (modified in place, must be float64 C-contiguous)."""
    ndim = len(ilower)
    X = _dv(X).reshape(-1, ndim)
    N = X.shape[0]
    if indices is None:
        indices = np.arange(N)
    indices = _iv(indices)
    if Xshift is None:
        Xshift = np.zeros((len(indices), ndim))
    Xshift = _dv(Xshift)
    V = _dv(V).reshape(N, depth)
    assert u.dtype == np.float64 and u.flags.c_contiguous
    lib().orc_spread(kernel_id(kernel), ndim, _d(_dv(dx)), _d(_dv(x_lower)), depth, _i(indices), _d(Xshift),
                     len(indices), _d(X), _d(V), _i(_iv(ilower)), _i(_iv(iupper)), _i(_iv(nugc)), _d(u))
    return u


def get_cell_index(X, x_lower, x_upper, dx, ilower, iupper):
    ndim = len(ilower)
    X2 = _dv(X).reshape(-1, ndim)
    out = np.zeros((X2.shape[0], ndim), dtype=np.int32)
    lib().orc_get_cell_indices(ndim, _d(X2), X2.shape[0], _d(_dv(x_lower)), _d(_dv(x_upper)), _d(_dv(dx)),
                               _i(_iv(ilower)), _i(_iv(iupper)), _i(out))
    return out


def build_local_indices(box_lo, box_hi, x_lower, x_upper, dx, ilower, iupper, X):
    ndim = len(ilower)
    X = _dv(X).reshape(-1, ndim)
    idx = np.zeros(X.shape[0], dtype=np.int32)
    n = lib().orc_build_local_indices(ndim, _i(_iv(box_lo)), _i(_iv(box_hi)), _d(_dv(x_lower)),
                                      _d(_dv(x_upper)), _d(_dv(dx)), _i(_iv(ilower)), _i(_iv(iupper)), _d(X),
                                      X.shape[0], _i(idx))
    return idx[:n].copy()


def side_axis_shape(ilower, iupper, gcw, axis):
    iu = list(iupper)
    iu[axis] += 1
    return ghosted_shape(ilower, iu, gcw)[1:]


def interp_side(kernel, dx, x_lower, ilower, iupper, gcw, q_axes, X, indices=None, Xshift=None):
    ndim = len(ilower)
    X = _dv(X).reshape(-1, ndim)
    N = X.shape[0]
    if indices is None:
        indices = np.arange(N)
    indices = _iv(indices)
    if Xshift is None:
        Xshift = np.zeros((len(indices), ndim))
    Xshift = _dv(Xshift)
    qs = [_dv(q) for q in q_axes]
    arr = (_dp * ndim)(*[_d(q) for q in qs])
    Q = np.full((N, ndim), np.nan)
    lib().orc_interp_side(kernel_id(kernel), ndim, _d(_dv(dx)), _d(_dv(x_lower)), _i(_iv(ilower)),
                          _i(_iv(iupper)), _i(_iv(gcw)), arr, _i(indices), _d(Xshift), len(indices), _d(X), _d(Q))
    return Q


def spread_side(kernel, dx, x_lower, ilower, iupper, gcw, q_axes, X, Q, indices=None, Xshift=None):
    """Accumulates into the float64 C-contiguous arrays q_axes (in place)."""
    ndim = len(ilower)
    X = _dv(X).reshape(-1, ndim)
    N = X.shape[0]
    if indices is None:
        indices = np.arange(N)
    indices = _iv(indices)
    if Xshift is None:
        Xshift = np.zeros((len(indices), ndim))
    Xshift = _dv(Xshift)
    for q in q_axes:
        assert q.dtype == np.float64 and q.flags.c_contiguous
    arr = (_dp * ndim)(*[_d(q) for q in q_axes])
    Q = _dv(Q).reshape(N, ndim)
    lib().orc_spread_side(kernel_id(kernel), ndim, _d(_dv(dx)), _d(_dv(x_lower)), _i(_iv(ilower)),
                          _i(_iv(iupper)), _i(_iv(gcw)), arr, _i(indices), _d(Xshift), len(indices), _d(X), _d(Q))
    return q_axes


def build_patch_entries(cell, ilower, iupper, idx_gcw, periodic_shift, domain_lower, dx, bin_size,
                        interior_only=False):
    ndim = len(ilower)
    cell = _iv(cell).reshape(-1, ndim)
    npts = cell.shape[0]
    cap = npts * (2 ** ndim) + 1
    idx = np.zeros(cap, dtype=np.int32)
    sh = np.zeros((cap, ndim))
    n = lib().orc_build_patch_entries(ndim, _i(cell), npts, _i(_iv(ilower)), _i(_iv(iupper)), int(idx_gcw),
                                      _i(_iv(periodic_shift)), _i(_iv(domain_lower)), _d(_dv(dx)),
                                      _i(_iv(bin_size)), int(bool(interior_only)), cap, _i(idx), _d(sh))
    return idx[:n].copy(), sh[:n].copy()


def bruteforce_spread_at(kernel, xg, dx, period, X, V):
    ndim = len(dx)
    X = _dv(X).reshape(-1, ndim)
    return lib().orc_bruteforce_spread_at(kernel_id(kernel), ndim, _d(_dv(xg)), _d(_dv(dx)), _d(_dv(period)),
                                          _d(X), _d(_dv(V)), X.shape[0])


def bin_ranks(cell, domain_lower, bin_size):
    """orc_bin_ranks: each point's position among its bin's points (cell's z
    plane inside the bin, then index); shared by all patches of a level."""
    cell = _iv(cell).reshape(-1, 3)
    out = np.zeros(max(1, cell.shape[0]), dtype=np.int64)
    lib().orc_bin_ranks(_i(cell), cell.shape[0], _i(_iv(domain_lower)), _i(_iv(bin_size)),
                        out.ctypes.data_as(_lp))
    return out


def order_entries_z(kernel, cell, indices, Xshift, X, x_lower, dx, domain_lower, bin_size, chunk=64, ranks=None):
    """Permutation of a canonical 3-D entry list into the GPU spread order for one
    array (bin/image group, 64-point chunk of the bin, stencil z-start, point).
    ranks: bin_ranks(cell, domain_lower, bin_size), computed here when None."""
    cell = _iv(cell).reshape(-1, 3)
    indices = _iv(indices)
    Xshift = _dv(Xshift).reshape(-1, 3)
    X = _dv(X).reshape(-1, 3)
    perm = np.zeros(len(indices), dtype=np.int32)
    if ranks is None:
        ranks = bin_ranks(cell, domain_lower, bin_size)
    lib().orc_order_entries_z(kernel_id(kernel), _i(cell), cell.shape[0], _i(indices), _d(Xshift), len(indices), _d(X),
                              _d(_dv(x_lower)), _d(_dv(dx)), _i(_iv(domain_lower)), _i(_iv(bin_size)), int(chunk),
                              np.ascontiguousarray(ranks, dtype=np.int64).ctypes.data_as(_lp), _i(perm))
    return perm


def spread_patch(kernel, centering_side, dx, x_lower, ilower, iupper, gcw, arrays, X, F, cells, idx_gcw,
                 periodic_shift, domain_lower, depth=1, bin_size=None, ranks=None):
    """LEInteractor::spread of one patch over its Lagrangian ghost box (entries
    with periodic images, LIndexSetData model) into `arrays` (side: one float64
    array per axis; cell: one (depth, ...) array), in the summation order the
    GPU library uses: canonical (bin, image, point) order in 2-D; in 3-D each
    array's entries re-ordered by orc_order_entries_z (64-point chunks of each
    bin, then stencil z-start)."""
    ndim = len(ilower)
    if bin_size is None:
        bin_size = [8] * ndim if ndim == 3 else [16] * ndim
    idx, sh = build_patch_entries(cells, ilower, iupper, idx_gcw, periodic_shift, domain_lower, dx, bin_size)
    if len(idx) == 0:
        return arrays
    X = _dv(X).reshape(-1, ndim)
    F = _dv(F).reshape(X.shape[0], -1)
    if ndim == 3 and ranks is None:
        ranks = bin_ranks(cells, domain_lower, bin_size)
    if centering_side:
        for a in range(ndim):
            xl = list(x_lower)
            xl[a] -= 0.5 * dx[a]
            iu = list(iupper)
            iu[a] += 1
            if ndim == 3:
                perm = order_entries_z(kernel, cells, idx, sh, X, xl, dx, domain_lower, bin_size, ranks=ranks)
                ii, ss = idx[perm], sh[perm]
            else:
                ii, ss = idx, sh
            q = arrays[a].reshape((1,) + arrays[a].shape)
            spread(kernel, dx, xl, ilower, iu, gcw, q, X, np.ascontiguousarray(F[:, a]), indices=ii, Xshift=ss)
    else:
        if ndim == 3:
            perm = order_entries_z(kernel, cells, idx, sh, X, x_lower, dx, domain_lower, bin_size, ranks=ranks)
            idx, sh = idx[perm], sh[perm]
        spread(kernel, dx, x_lower, ilower, iupper, gcw, arrays, X, F, indices=idx, Xshift=sh, depth=depth)
    return arrays


def lagrangian_force(ndim, X, F=None, U=None, springs=None, beams=None, targets=None, spring_laws=None):
    """IBStandardForceGen forces, reference scatter order (orc_lagrangian_force).
    springs = (master, slave, kappa, rest); beams = (prev, master, next,
    rigidity, curvature); targets = (node, kappa, eta, X0).  Returns F + forces.
    spring_laws = (law, offsets, params): a law (0 Hookean, 1 linear strain,
    2 polynomial) and params[offsets[k]:offsets[k+1]] per spring, in place of
    the Hookean kappa and rest (orc_lagrangian_force_laws)."""
    X = _dv(X).reshape(-1, ndim)
    n = X.shape[0]
    out = np.zeros((n, ndim)) if F is None else _dv(F).reshape(n, ndim).copy()
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    if springs is not None:
        sm, ss, sk, sr = i32(springs[0]), i32(springs[1]), _dv(springs[2]), _dv(springs[3])
    else:
        sm = ss = np.zeros(1, np.int32)
        sk = sr = np.zeros(1)
    if beams is not None:
        bp, bm, bn, bK, bc = i32(beams[0]), i32(beams[1]), i32(beams[2]), _dv(beams[3]), _dv(beams[4])
    else:
        bp = bm = bn = np.zeros(1, np.int32)
        bK = bc = np.zeros(max(1, ndim))
    if targets is not None:
        tn, tK, tE, tX0 = i32(targets[0]), _dv(targets[1]), _dv(targets[2]), _dv(targets[3])
    else:
        tn = np.zeros(1, np.int32)
        tK = tE = tX0 = np.zeros(max(1, ndim))
    Ud = None if U is None else _dv(U).reshape(n, ndim)
    nsp = len(springs[0]) if springs is not None else 0
    nbm = len(beams[0]) if beams is not None else 0
    ntg = len(targets[0]) if targets is not None else 0
    Up = _d(Ud) if Ud is not None else None
    if spring_laws is None:
        lib().orc_lagrangian_force(ndim, n, _d(out), _d(X), Up, nsp, _i(sm), _i(ss), _d(sk), _d(sr),
                                   nbm, _i(bm), _i(bn), _i(bp), _d(bK), _d(bc), ntg, _i(tn), _d(tK), _d(tE), _d(tX0))
        return out
    law, poff, par = i32(spring_laws[0]), i32(spring_laws[1]), _dv(spring_laws[2])
    assert len(law) == nsp and len(poff) == nsp + 1
    L = lib()
    if not getattr(L, "_lfl_types", False):
        L.orc_lagrangian_force_laws.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 3 + \
            [ctypes.c_int] + [ctypes.c_void_p] * 5 + [ctypes.c_int] + [ctypes.c_void_p] * 5 + \
            [ctypes.c_int] + [ctypes.c_void_p] * 4
        L._lfl_types = True
    law = law if nsp else np.zeros(1, np.int32)
    par = par if len(par) else np.zeros(1)
    L.orc_lagrangian_force_laws(ndim, n, _d(out), _d(X), Up, nsp, _i(sm), _i(ss), _i(law), _i(poff), _d(par),
                                nbm, _i(bm), _i(bn), _i(bp), _d(bK), _d(bc), ntg, _i(tn), _d(tK), _d(tE), _d(tX0))
    return out


def redistribute(X, x_lower, x_upper, dx, dom_lo, dom_hi, periodic, boxes, lag=None, error_if_leave=False,
                 max_wraps=64):
    """LDataManager redistribution of one level (orc_redistribute).  boxes:
    list of (lower, upper) patch boxes in level order.  Returns (status, X
    moved into the domain, new_order, patch_offsets, periodic offsets)."""
    ndim = len(dom_lo)
    X = _dv(X).reshape(-1, ndim).copy()
    n = X.shape[0]
    plo = np.ascontiguousarray([b[0] for b in boxes], dtype=np.int32).reshape(-1)
    phi = np.ascontiguousarray([b[1] for b in boxes], dtype=np.int32).reshape(-1)
    order = np.zeros(max(n, 1), dtype=np.int32)
    poff = np.zeros(len(boxes) + 1, dtype=np.int32)
    offs = np.zeros((max(n, 1), ndim), dtype=np.int32)
    lg = None if lag is None else np.ascontiguousarray(lag, dtype=np.int32)
    L = lib()
    if not getattr(L, "_redist_types", False):
        L.orc_redistribute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_int)] + [ctypes.POINTER(ctypes.c_double)] * 3 + \
            [ctypes.POINTER(ctypes.c_int)] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
            [ctypes.POINTER(ctypes.c_int)] * 5
        L._redist_types = True
    rc = L.orc_redistribute(ndim, n, _d(X), _i(lg) if lg is not None else None, _d(_dv(x_lower)), _d(_dv(x_upper)),
                            _d(_dv(dx)), _i(_iv(dom_lo)), _i(_iv(dom_hi)), _i(_iv([1 if p else 0 for p in periodic])),
                            1 if error_if_leave else 0, max_wraps, len(boxes), _i(plo), _i(phi), _i(order), _i(poff),
                            _i(offs))
    return rc, X, order[:n], poff, offs[:n]


def fe_quadrature_points(elem_nodes, phi, JxW, X_nodes, F_nodes=None):
    """FEDataManager::spread's point source (orc_fe_quadrature_points)."""
    en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    nelem, nbasis = en.shape
    phi = _dv(phi)
    nqp = phi.shape[1]
    Xn = _dv(X_nodes)
    ndim = Xn.shape[1]
    Xq = np.zeros((nelem * nqp, ndim))
    Fn = None if F_nodes is None else _dv(F_nodes)
    nvars = 0 if Fn is None else Fn.shape[1]
    Fq = np.zeros((nelem * nqp, max(nvars, 1)))
    L = lib()
    if not getattr(L, "_feq_types", False):
        L.orc_fe_quadrature_points.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)] + \
            [ctypes.POINTER(ctypes.c_double)] * 6
        L._feq_types = True
    Jw = None if JxW is None else _dv(JxW)
    L.orc_fe_quadrature_points(ndim, nvars, nelem, nbasis, nqp, _i(en), _d(phi), _d(Jw) if Jw is not None else None,
                               _d(Xn), _d(Fn) if Fn is not None else None, _d(Xq), _d(Fq))
    return Xq, (Fq[:, :nvars] if Fn is not None else None)


def fe_integrate_rhs(elem_nodes, phi, JxW, U_qp, nnodes, F=None):
    """FEDataManager::interpWeighted's right-hand side (orc_fe_integrate_rhs); returns F + rhs."""
    en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    nelem, nbasis = en.shape
    phi = _dv(phi)
    nqp = phi.shape[1]
    U = _dv(U_qp)
    nvars = U.shape[1]
    out = np.zeros((nnodes, nvars)) if F is None else _dv(F).copy()
    L = lib()
    if not getattr(L, "_ferhs_types", False):
        L.orc_fe_integrate_rhs.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)] + \
            [ctypes.POINTER(ctypes.c_double)] * 4
        L._ferhs_types = True
    L.orc_fe_integrate_rhs(nvars, nelem, nbasis, nqp, _i(en), _d(phi), _d(_dv(JxW)), _d(U), _d(out))
    return out


FE_ELEM_TYPES = ["EDGE2", "EDGE3", "TRI3", "TRI6", "QUAD4", "QUAD8", "QUAD9", "TET4", "TET10", "HEX8", "HEX20",
                 "HEX27"]


def fe_nodal_mass(elem_nodes, phi, JxW, nnodes):
    """Nodal quadrature's diagonal mass (orc_fe_nodal_mass): phi [nbasis][nqp]
    and JxW [nelem][nqp] of the nodal rule."""
    en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    nelem, nbasis = en.shape
    phi = _dv(phi)
    M = np.zeros(max(nnodes, 1))
    L = lib()
    L.orc_fe_nodal_mass.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int), ctypes.c_int] + \
        [ctypes.POINTER(ctypes.c_double)] * 3
    L.orc_fe_nodal_mass(nelem, nbasis, nnodes, _i(en), phi.shape[1], _d(phi), _d(_dv(JxW)), _d(M))
    return M[:nnodes]


def fe_adaptive_orders(elem_type, elem_nodes, X_nodes, quad_type, point_density, dx_min):
    """getQuadratureKey's adaptive order per element (orc_fe_adaptive_orders);
    quad_type 0 = QGAUSS, 1 = QGRID."""
    en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    nelem, nbasis = en.shape
    X = _dv(X_nodes)
    out = np.zeros(max(nelem, 1), dtype=np.int32)
    L = lib()
    L.orc_fe_adaptive_orders.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int),
                                                              ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                                              ctypes.c_double, ctypes.c_double,
                                                              ctypes.POINTER(ctypes.c_int)]
    L.orc_fe_adaptive_orders(FE_ELEM_TYPES.index(elem_type), X.shape[1], nelem, nbasis, _i(en), _d(X), quad_type,
                             point_density, dx_min, _i(out))
    return out[:nelem]


def fe_max_edge_length(elem_type, nodes, X_nodes):
    X = _dv(X_nodes)
    n = np.ascontiguousarray(nodes, dtype=np.int32)
    L = lib()
    L.orc_fe_max_edge_length.restype = ctypes.c_double
    L.orc_fe_max_edge_length.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_double)]
    return L.orc_fe_max_edge_length(FE_ELEM_TYPES.index(elem_type), X.shape[1], _i(n), _d(X))


def fe_rules_layout(phis, elem_rule):
    """(rule_nqp, rule_phi_off, concatenated phi, qp_off) for tables phis[r]
    ([nbasis][nqp_r]) and a rule index per element."""
    nqp = np.array([np.shape(p)[1] for p in phis], dtype=np.int32)
    phi_off = np.zeros(len(phis), dtype=np.int64)
    phi_off[1:] = np.cumsum([np.size(p) for p in phis])[:-1]
    phi = np.concatenate([_dv(p).reshape(-1) for p in phis])
    er = np.ascontiguousarray(elem_rule, dtype=np.int32)
    qp_off = np.zeros(len(er) + 1, dtype=np.int64)
    qp_off[1:] = np.cumsum(nqp[er])
    return nqp, phi_off, phi, qp_off


def fe_quadrature_points_rules(elem_nodes, phis, elem_rule, JxW, X_nodes, F_nodes=None):
    """The point source with a rule per element (orc_fe_quadrature_points_rules);
    JxW flat, qp_off[e] onwards for element e."""
    en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    nelem, nbasis = en.shape
    nqp, phi_off, phi, qp_off = fe_rules_layout(phis, elem_rule)
    er = np.ascontiguousarray(elem_rule, dtype=np.int32)
    Xn = _dv(X_nodes)
    ndim = Xn.shape[1]
    tot = int(qp_off[-1])
    Xq = np.zeros((max(tot, 1), ndim))
    Fn = None if F_nodes is None else _dv(F_nodes)
    nvars = 0 if Fn is None else Fn.shape[1]
    Fq = np.zeros((max(tot, 1), max(nvars, 1)))
    Jw = None if JxW is None else _dv(JxW)
    L = lib()
    vp = ctypes.c_void_p
    L.orc_fe_quadrature_points_rules.argtypes = [ctypes.c_int] * 4 + [vp] * 10
    L.orc_fe_quadrature_points_rules(ndim, nvars, nelem, nbasis, _i(en), _i(er), _i(nqp), phi_off.ctypes.data,
                                     _d(phi), qp_off.ctypes.data, _d(Jw) if Jw is not None else None, _d(Xn),
                                     _d(Fn) if Fn is not None else None, _d(Xq), _d(Fq))
    return Xq[:tot], (Fq[:tot, :nvars] if Fn is not None else None)


def fe_integrate_rhs_rules(elem_nodes, phis, elem_rule, JxW, U_qp, nnodes, F=None):
    """The weighted right-hand side with a rule per element (orc_fe_integrate_rhs_rules)."""
    en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
    nelem, nbasis = en.shape
    nqp, phi_off, phi, qp_off = fe_rules_layout(phis, elem_rule)
    er = np.ascontiguousarray(elem_rule, dtype=np.int32)
    U = _dv(U_qp)
    nvars = U.shape[1]
    out = np.zeros((nnodes, nvars)) if F is None else _dv(F).copy()
    L = lib()
    vp = ctypes.c_void_p
    L.orc_fe_integrate_rhs_rules.argtypes = [ctypes.c_int] * 3 + [vp] * 9
    L.orc_fe_integrate_rhs_rules(nvars, nelem, nbasis, _i(en), _i(er), _i(nqp), phi_off.ctypes.data, _d(phi),
                                 qp_off.ctypes.data, _d(_dv(JxW)), _d(U), _d(out))
    return out


def phys_bdry(arrays, ilower, iupper, gcw, dx, boxes, abg, side, adjoint, depth=1):
    """CartCell/CartSideRobinPhysBdryOp LINEAR ops on one patch (orc_phys_bdry),
    in place on float64 C-contiguous numpy arrays (cell: [arr]; side: one per axis).
    boxes: list of (codim, location, lo, hi); abg: [nboxes][narrays][3]."""
    ndim = len(ilower)
    for a in arrays:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    bx = np.zeros((max(len(boxes), 1), 8), dtype=np.int32)
    for k, (c, l, lo, hi) in enumerate(boxes):
        bx[k, 0], bx[k, 1] = c, l
        bx[k, 2:2 + ndim] = lo
        bx[k, 5:5 + ndim] = hi
    ab = _dv(abg) if len(boxes) else np.zeros(3)
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrays))(*[_d(a) for a in arrays])
    L = lib()
    if not getattr(L, "_pb_types", False):
        L.orc_phys_bdry.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)] * 2 + \
            [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.c_int,
             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L._pb_types = True
    L.orc_phys_bdry(ndim, 1 if side else 0, depth, gcw, _i(_iv(ilower)), _i(_iv(iupper)), _d(_dv(dx)), ptrs,
                    len(boxes), _i(bx), _d(ab), 1 if adjoint else 0)
    return arrays


def phys_bdry_ncoef(ndim, box, j, side):
    """Coefficient values of a codim-1 box (codim, location, lo, hi) for array j."""
    _, loc, lo, hi = box
    m = 1
    for d in range(ndim):
        if d != loc // 2:
            m *= hi[d] - lo[d] + 1 + (1 if side and d == j else 0)
    return m


def phys_bdry_coefs(arrays, ilower, iupper, gcw, dx, boxes, coefs, side, adjoint, depth=1):
    """orc_phys_bdry_coefs: the operators with Robin coefficients varying over
    each codim-1 box.  coefs: for every codim-1 box in order and every array,
    a, b, g over the box's tangential cells (phys_bdry_ncoef values each, the
    lower tangential axis fastest), concatenated."""
    ndim = len(ilower)
    for a in arrays:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    bx = np.zeros((max(len(boxes), 1), 8), dtype=np.int32)
    for k, (c, l, lo, hi) in enumerate(boxes):
        bx[k, 0], bx[k, 1] = c, l
        bx[k, 2:2 + ndim] = lo
        bx[k, 5:5 + ndim] = hi
    cf = _dv(coefs) if len(coefs) else np.zeros(3)
    L = lib()
    if not getattr(L, "_pbc_types", False):
        L.orc_phys_bdry_ncoefs.restype = ctypes.c_long
        L.orc_phys_bdry_ncoefs.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)]
        L.orc_phys_bdry_coefs.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_int)] * 2 + \
            [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.c_int,
             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L._pbc_types = True
    need = L.orc_phys_bdry_ncoefs(ndim, 1 if side else 0, depth, len(boxes), _i(bx))
    if len(coefs) != need:
        raise ValueError("coefs holds %d values, the boxes need %d" % (len(coefs), need))
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrays))(*[_d(a) for a in arrays])
    L.orc_phys_bdry_coefs(ndim, 1 if side else 0, depth, gcw, _i(_iv(ilower)), _i(_iv(iupper)), _d(_dv(dx)), ptrs,
                          len(boxes), _i(bx), _d(cf), 1 if adjoint else 0)
    return arrays
