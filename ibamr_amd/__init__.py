"""ibamr_amd -- MI355X (gfx950) Eulerian-Lagrangian coupling for IBAMR-style
immersed-boundary codes: regularized-delta spreading and interpolation
(IB_4, IB_6, BSPLINE_3..6, PIECEWISE_LINEAR) on cell- and side-centred patch data.

Python host layer over the C ABI of libibamr_amd.so (include/ibamr_amd.h).
Device memory comes from torch tensors (plumbing only); every computation runs
in the HIP library.  Importing works without a GPU; the library must have been
built (python -m ibamr_amd.build), otherwise loading raises.

Interface mirror of the reference (IBTK):
  is_known_kernel / get_stencil_size / get_minimum_ghost_width
      LEInteractor::isKnownKernel / getStencilSize / getMinimumGhostWidth
      (ibtk/src/lagrangian/LEInteractor.cpp:2035-2114)
  LevelPlan.spread / LevelPlan.interp
      LDataManager::spread / LDataManager::interp for one level
      (ibtk/src/lagrangian/LDataManager.cpp:552-667, 699-813), which call
      LEInteractor::spread / interpolate per patch.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# IBAMR_AMD_LIB: another build of the library (a different path), e.g. to compare builds
LIB_PATH = os.environ.get("IBAMR_AMD_LIB") or os.path.join(_HERE, "lib", "libibamr_amd.so")

IBAMR_OK = 0
IBAMR_ERR_BAD_KERNEL = -1
IBAMR_ERR_GHOST_WIDTH = -2
IBAMR_ERR_DEPTH = -3
IBAMR_ERR_INVALID = -4
IBAMR_ERR_OUT_OF_DOMAIN = -5
IBAMR_ERR_CAPACITY = -6
IBAMR_ERR_DEVICE = -7
IBAMR_ERR_NOT_BINNED = -8

CELL, SIDE = 0, 1
FP64, FP32C = 0, 1
ACCUMULATE, OVERWRITE, ACCUMULATE_INTERIOR = 0, 1, 2

# Every symbol include/ibamr_amd.h declares (checked by the CPU test suite).
C_API = [
    "ibamr_is_known_kernel", "ibamr_get_stencil_size", "ibamr_get_minimum_ghost_width",
    "ibamr_plan_create", "ibamr_plan_create_patch", "ibamr_plan_destroy", "ibamr_plan_bind_data", "ibamr_plan_bin",
    "ibamr_plan_bin_async", "ibamr_plan_check", "ibamr_plan_spread", "ibamr_plan_spread_ds", "ibamr_plan_interp",
    "ibamr_plan_fill_periodic_ghosts", "ibamr_plan_fill_ghosts", "ibamr_forward_euler_step", "ibamr_forward_euler_step_periodic", "ibamr_plan_set_timing", "ibamr_plan_kernel_times",
    "ibamr_plan_info", "ibamr_plan_set_tile_height", "ibamr_plan_get_tile_height", "ibamr_plan_autotune",
    "ibamr_plan_set_tile_shape", "ibamr_plan_get_tile_shape", "ibamr_plan_get_cell_indices",
    "ibamr_plan_redistribute", "ibamr_permute_rows",
    "ibamr_plan_set_physical_boundary", "ibamr_plan_set_physical_boundary_coefs",
    "ibamr_plan_accumulate_physical_boundary", "ibamr_plan_fill_physical_boundary", "ibamr_fe_quadrature_points",
    "ibamr_fe_mesh_create", "ibamr_fe_integrate_rhs", "ibamr_fe_mesh_destroy",
    "ibamr_fe_adaptive_quadrature_orders", "ibamr_fe_quadrature_points_rules", "ibamr_fe_integrate_rhs_rules",
    "ibamr_fe_nodal_mass", "ibamr_fe_nodal_weight",
    "ibamr_last_error", "ibamr_forces_create", "ibamr_forces_compute", "ibamr_forces_destroy",
    "ibamr_forces_register_spring_function", "ibamr_forces_set_spring_parameters",
    "ibamr_s0_interp", "ibamr_s0_spread", "ibamr_s0_interp_noalloc",
] + ["lagrangian_%s_%s%s_" % (k, op, d) for k in ("ib_4", "ib_6", "bspline_3", "piecewise_linear",
                                                        "bspline_4", "bspline_5", "bspline_6")
     for op in ("interp", "spread") for d in ("2d", "3d")]


class IBAMRError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ibamr_amd error %d: %s" % (code, msg))
        self.code = code


class PatchGeom(ctypes.Structure):
    _fields_ = [("ilower", ctypes.c_int * 3), ("iupper", ctypes.c_int * 3),
                ("x_lower", ctypes.c_double * 3), ("x_upper", ctypes.c_double * 3)]


class LevelGeom(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int), ("dx", ctypes.c_double * 3),
                ("domain_lower", ctypes.c_int * 3), ("domain_upper", ctypes.c_int * 3),
                ("domain_x_lower", ctypes.c_double * 3), ("domain_x_upper", ctypes.c_double * 3),
                ("periodic_shift", ctypes.c_int * 3), ("npatches", ctypes.c_int),
                ("patches", ctypes.POINTER(PatchGeom))]


class BdryBox(ctypes.Structure):
    _fields_ = [("patch", ctypes.c_int), ("codim", ctypes.c_int), ("location", ctypes.c_int),
                ("lower", ctypes.c_int * 3), ("upper", ctypes.c_int * 3)]


class PlanConfig(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_char_p), ("centering", ctypes.c_int), ("precision", ctypes.c_int),
                ("depth", ctypes.c_int), ("data_gcw", ctypes.c_int * 3), ("idx_gcw", ctypes.c_int),
                ("max_points", ctypes.c_int), ("drift_cells", ctypes.c_int), ("tile", ctypes.c_int)]


_LIB = None


def lib():
    """Loads libibamr_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError("libibamr_amd.so is not built (%s); run python -m ibamr_amd.build" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.ibamr_last_error.restype = ctypes.c_char_p
    L.ibamr_is_known_kernel.argtypes = [ctypes.c_char_p]
    L.ibamr_get_stencil_size.argtypes = [ctypes.c_char_p]
    L.ibamr_get_minimum_ghost_width.argtypes = [ctypes.c_char_p]
    L.ibamr_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(LevelGeom), ctypes.POINTER(PlanConfig)]
    L.ibamr_plan_create_patch.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(PatchGeom), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(PlanConfig)]
    L.ibamr_plan_destroy.argtypes = [vp]
    L.ibamr_plan_set_timing.argtypes = [vp, ctypes.c_int]
    ip_, dp_ = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.ibamr_forces_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int, ip_, ip_, dp_, dp_,
                                      ctypes.c_int, ip_, ip_, ip_, dp_, dp_, ctypes.c_int, ip_, dp_, dp_, dp_]
    L.ibamr_forces_compute.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    L.ibamr_forces_destroy.argtypes = [vp]
    L.ibamr_forces_register_spring_function.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.ibamr_forces_set_spring_parameters.argtypes = [vp, ip_, ip_, dp_]
    L.ibamr_forward_euler_step.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_long, vp]
    L.ibamr_forward_euler_step_periodic.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_long,
                                                    ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), vp]
    L.ibamr_plan_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    L.ibamr_plan_bind_data.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.ibamr_plan_bin.argtypes = [vp, vp, ctypes.c_int, vp]
    L.ibamr_plan_bin_async.argtypes = [vp, vp, ctypes.c_int, vp]
    L.ibamr_plan_check.argtypes = [vp, vp]
    L.ibamr_plan_spread.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp]
    L.ibamr_plan_spread_ds.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
    L.ibamr_plan_interp.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    L.ibamr_plan_fill_periodic_ghosts.argtypes = [vp, ctypes.c_int, vp]
    L.ibamr_plan_fill_ghosts.argtypes = [vp, ctypes.c_int, vp]
    L.ibamr_plan_redistribute.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
    L.ibamr_permute_rows.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    L.ibamr_plan_set_physical_boundary.argtypes = [vp, ctypes.POINTER(BdryBox), ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_double)]
    L.ibamr_plan_set_physical_boundary_coefs.argtypes = [vp, ctypes.POINTER(BdryBox), ctypes.c_int,
                                                         ctypes.POINTER(ctypes.c_double), ctypes.c_long]
    L.ibamr_plan_accumulate_physical_boundary.argtypes = [vp, ctypes.c_int, vp]
    L.ibamr_plan_fill_physical_boundary.argtypes = [vp, ctypes.c_int, vp]
    L.ibamr_fe_quadrature_points.argtypes = [ctypes.c_int] * 6 + [vp] * 7 + [ctypes.c_int, vp]
    L.ibamr_fe_mesh_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int)]
    L.ibamr_fe_integrate_rhs.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, vp]
    L.ibamr_fe_mesh_destroy.argtypes = [vp]
    L.ibamr_fe_adaptive_quadrature_orders.argtypes = [ctypes.c_int] * 5 + [vp, vp, ctypes.c_int, ctypes.c_double,
                                                                          ctypes.c_double, vp, vp]
    L.ibamr_fe_quadrature_points_rules.argtypes = [ctypes.c_int] * 5 + [vp, ctypes.c_int] + [vp] * 5 + \
        [ctypes.c_long] + [vp] * 5 + [ctypes.c_int, vp]
    L.ibamr_fe_integrate_rhs_rules.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [vp] * 7 + [ctypes.c_int, vp,
                                                                                         ctypes.c_int, vp]
    L.ibamr_fe_nodal_mass.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
    L.ibamr_fe_nodal_weight.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int,
                                        ctypes.c_int, vp]
    L.ibamr_plan_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)]
    L.ibamr_plan_set_tile_height.argtypes = [vp, ctypes.c_int]
    L.ibamr_plan_get_tile_height.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.ibamr_plan_set_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.ibamr_plan_get_cell_indices.argtypes = [vp, vp, vp]
    L.ibamr_plan_get_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.ibamr_plan_autotune.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int), vp]
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.ibamr_s0_interp.argtypes = [ctypes.c_char_p, ctypes.c_int, dp, dp, ctypes.c_int, ip, ip, ip, dp, ip, dp,
                                  ctypes.c_int, dp, dp]
    L.ibamr_s0_spread.argtypes = [ctypes.c_char_p, ctypes.c_int, dp, dp, ctypes.c_int, ip, dp, ctypes.c_int, dp,
                                  dp, ip, ip, ip, dp]
    _LIB = L
    return L


def _check(rc):
    if rc != IBAMR_OK:
        raise IBAMRError(rc, lib().ibamr_last_error().decode())
    return rc


def is_known_kernel(name):
    return bool(lib().ibamr_is_known_kernel(name.encode()))


def get_stencil_size(name):
    v = lib().ibamr_get_stencil_size(name.encode())
    if v < 0:
        _check(v)
    return v


def get_minimum_ghost_width(name):
    v = lib().ibamr_get_minimum_ghost_width(name.encode())
    if v < 0:
        _check(v)
    return v


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Patch:
    """A patch box on a level (cell indices) and its physical extent."""

    def __init__(self, ilower, iupper, x_lower, x_upper):
        self.ilower, self.iupper = list(ilower), list(iupper)
        self.x_lower, self.x_upper = list(x_lower), list(x_upper)


class Level:
    """Level geometry: dx, domain box and extent, periodic shift, patches."""

    def __init__(self, dx, domain_lower, domain_upper, domain_x_lower, domain_x_upper, periodic_shift, patches):
        self.ndim = len(dx)
        self.dx = list(dx)
        self.domain_lower, self.domain_upper = list(domain_lower), list(domain_upper)
        self.domain_x_lower, self.domain_x_upper = list(domain_x_lower), list(domain_x_upper)
        self.periodic_shift = list(periodic_shift)
        self.patches = list(patches)

    @staticmethod
    def uniform_periodic(n, ndim=3, length=1.0):
        """One patch covering a periodic box [0, length)^ndim with n cells per axis."""
        h = length / n
        p = Patch([0] * ndim, [n - 1] * ndim, [0.0] * ndim, [length] * ndim)
        return Level([h] * ndim, [0] * ndim, [n - 1] * ndim, [0.0] * ndim, [length] * ndim, [n] * ndim, [p])

    def to_c(self):
        pg = (PatchGeom * len(self.patches))()
        for i, p in enumerate(self.patches):
            for d in range(self.ndim):
                pg[i].ilower[d] = p.ilower[d]
                pg[i].iupper[d] = p.iupper[d]
                pg[i].x_lower[d] = p.x_lower[d]
                pg[i].x_upper[d] = p.x_upper[d]
        L = LevelGeom()
        L.ndim = self.ndim
        for d in range(self.ndim):
            L.dx[d] = self.dx[d]
            L.domain_lower[d] = self.domain_lower[d]
            L.domain_upper[d] = self.domain_upper[d]
            L.domain_x_lower[d] = self.domain_x_lower[d]
            L.domain_x_upper[d] = self.domain_x_upper[d]
            L.periodic_shift[d] = self.periodic_shift[d]
        L.npatches = len(self.patches)
        L.patches = ctypes.cast(pg, ctypes.POINTER(PatchGeom))
        return L, pg


def array_shape(level, patch_index, gcw, centering, axis=0, depth=1):
    """Shape (torch/numpy, C order) of one ghosted Eulerian array of a patch:
    (depth, n2, n1, n0) for cell data, (n2, n1, n0) of axis `axis` for side data."""
    p = level.patches[patch_index]
    n = [p.iupper[d] - p.ilower[d] + 1 + 2 * gcw[d] + (1 if (centering == SIDE and d == axis) else 0)
         for d in range(level.ndim)]
    shp = tuple(reversed(n))
    return shp if centering == SIDE else (depth,) + shp


class LevelPlan:
    """Spread/interp plan for one level (the LDataManager::spread/interp level loop).

    With box=(lower, upper) the level must hold one patch, and the plan is the
    raw-array LEInteractor form instead (ibamr_plan_create_patch): no periodic
    images, points taken when their cell (getCellIndex on the patch) lies in box."""

    def __init__(self, level, kernel="IB_4", centering=SIDE, precision=FP64, depth=1, data_gcw=None,
                 idx_gcw=0, max_points=0, drift_cells=0, box=None):
        self.level = level
        self.kernel = kernel
        self.centering = centering
        self.precision = precision
        self.depth = depth
        g = get_minimum_ghost_width(kernel) if is_known_kernel(kernel) else 1
        self.data_gcw = list(data_gcw) if data_gcw is not None else [g] * level.ndim
        Lc, self._pg = level.to_c()
        cfg = PlanConfig()
        cfg.kernel = kernel.encode()
        cfg.centering = centering
        cfg.precision = precision
        cfg.depth = depth
        for d in range(level.ndim):
            cfg.data_gcw[d] = self.data_gcw[d]
        cfg.idx_gcw = idx_gcw
        cfg.max_points = max_points
        cfg.drift_cells = drift_cells
        cfg.tile = 0
        h = ctypes.c_void_p()
        if box is None:
            _check(lib().ibamr_plan_create(ctypes.byref(h), ctypes.byref(Lc), ctypes.byref(cfg)))
        else:
            assert len(level.patches) == 1, "a box plan covers one patch"
            lo = (ctypes.c_int * 3)(*(list(box[0]) + [0] * (3 - level.ndim)))
            hi = (ctypes.c_int * 3)(*(list(box[1]) + [0] * (3 - level.ndim)))
            dx = (ctypes.c_double * 3)(*(level.dx + [1.0] * (3 - level.ndim)))
            _check(lib().ibamr_plan_create_patch(ctypes.byref(h), level.ndim, dx, self._pg, lo, hi,
                                                 ctypes.byref(cfg)))
        self._h = h
        self._bound = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().ibamr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def narr(self):
        return self.level.ndim if self.centering == SIDE else 1

    def info(self):
        nt, ts, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().ibamr_plan_info(self._h, ctypes.byref(nt), ctypes.byref(ts), ctypes.byref(nb)))
        return {"ntiles": nt.value, "tile": ts.value, "nbins": nb.value}

    @property
    def tile_height(self):
        """z extent of the spread tile shape in use (3-D)."""
        h = ctypes.c_int()
        _check(lib().ibamr_plan_get_tile_height(self._h, ctypes.byref(h)))
        return h.value

    @tile_height.setter
    def tile_height(self, h):
        _check(lib().ibamr_plan_set_tile_height(self._h, int(h)))

    @property
    def tile_shape(self):
        """(x, y, z) extents of the spread tile shape in use."""
        sh = (ctypes.c_int * 3)()
        _check(lib().ibamr_plan_get_tile_shape(self._h, sh))
        return tuple(sh)

    @tile_shape.setter
    def tile_shape(self, shape):
        _check(lib().ibamr_plan_set_tile_shape(self._h, (ctypes.c_int * 3)(*[int(v) for v in shape])))

    def autotune(self, slot, F, X, reps=3, stream=None):
        """Time the spread of F at X (binned) into the arrays bound to `slot`
        (overwritten) with every built tile shape and keep the fastest.
        Returns {(x, y, z): ms per spread call}."""
        n = 8  # IBAMR_MAX_TILE_SHAPES
        sh, ms, nc = (ctypes.c_int * (3 * n))(), (ctypes.c_float * n)(), ctypes.c_int()
        _check(lib().ibamr_plan_autotune(self._h, slot, _ptr(F), _ptr(X), reps, sh, ms, ctypes.byref(nc),
                                         _stream_ptr(stream)))
        return {(sh[3 * i], sh[3 * i + 1], sh[3 * i + 2]): ms[i] for i in range(nc.value)}

    def bind(self, slot, arrays):
        """arrays: flat list [patch0 axis0, patch0 axis1, ..., patch1 axis0, ...] of device tensors."""
        ptrs = (ctypes.c_void_p * len(arrays))(*[a.data_ptr() for a in arrays])
        _check(lib().ibamr_plan_bind_data(self._h, slot, ptrs))
        self._bound[slot] = list(arrays)  # keep alive

    def bin(self, X, stream=None, check=True):
        n = X.shape[0]
        fn = lib().ibamr_plan_bin if check else lib().ibamr_plan_bin_async
        _check(fn(self._h, _ptr(X), n, _stream_ptr(stream)))

    def check(self, stream=None):
        _check(lib().ibamr_plan_check(self._h, _stream_ptr(stream)))

    def cell_indices(self, n, stream=None):
        """getCellIndex of each of the n points of the last binning, as the
        binning computed it: device int32 [n][ndim] (INT_MIN: does not fit)."""
        import torch
        out = torch.empty((n, self.level.ndim), dtype=torch.int32, device="cuda")
        _check(lib().ibamr_plan_get_cell_indices(self._h, _ptr(out), _stream_ptr(stream)))
        return out

    def spread(self, slot, F, X, mode=ACCUMULATE, stream=None, ds=None):
        """LDataManager::spread into the arrays bound to slot; with ds (device
        [n] of F's type) the points carry F*ds (the ds_data overload)."""
        if ds is None:
            _check(lib().ibamr_plan_spread(self._h, slot, _ptr(F), _ptr(X), mode, _stream_ptr(stream)))
        else:
            if ds.dtype != F.dtype or ds.numel() != F.shape[0]:
                raise ValueError("ds must hold one value of F's type per point")
            _check(lib().ibamr_plan_spread_ds(self._h, slot, _ptr(F), _ptr(X), _ptr(ds.contiguous()), mode,
                                              _stream_ptr(stream)))

    def interp(self, slot, U, X, stream=None):
        _check(lib().ibamr_plan_interp(self._h, slot, _ptr(U), _ptr(X), _stream_ptr(stream)))

    def fill_periodic_ghosts(self, slot, stream=None):
        _check(lib().ibamr_plan_fill_periodic_ghosts(self._h, slot, _stream_ptr(stream)))

    def redistribute(self, X, lag_index=None, error_if_points_leave=False, periodic_offsets=False, stream=None):
        """LDataManager::begin/endDataRedistribution for the points X (device
        float64 [n][ndim], moved into the domain in place).  Returns
        (new_order [n] int32, patch_offsets [npatches+1] int32, offsets
        [n][ndim] int32 or None), all device tensors; new_order[k] is the old
        position of the point that becomes k.  Bin again after permuting."""
        import torch
        n = X.shape[0] if X.dim() > 0 else 0
        order = torch.empty(max(n, 1), dtype=torch.int32, device=X.device)
        poff = torch.empty(len(self.level.patches) + 1, dtype=torch.int32, device=X.device)
        offs = torch.empty((max(n, 1), self.level.ndim), dtype=torch.int32, device=X.device) if periodic_offsets \
            else None
        lag = None if lag_index is None else lag_index.to(torch.int32).contiguous()
        _check(lib().ibamr_plan_redistribute(self._h, _ptr(X), _ptr(lag), n, 1 if error_if_points_leave else 0,
                                             _ptr(order), _ptr(poff), _ptr(offs), _stream_ptr(stream)))
        return order[:n], poff, (offs[:n] if offs is not None else None)

    @staticmethod
    def _bdry_boxes(boxes):
        arr = (BdryBox * max(1, len(boxes)))()
        for k, (p, c, l, lo, hi) in enumerate(boxes):
            arr[k].patch, arr[k].codim, arr[k].location = p, c, l
            for d in range(len(lo)):
                arr[k].lower[d], arr[k].upper[d] = lo[d], hi[d]
        return arr

    def set_physical_boundary(self, boxes, abg):
        """boxes: [(patch, codim, location, lower, upper)] in each patch's order;
        abg: [nboxes][arrays][3] Robin coefficients (a, b, g) of codim-1 boxes."""
        import numpy as np
        arr = self._bdry_boxes(boxes)
        ab = np.ascontiguousarray(abg, dtype=np.float64).reshape(-1)
        if ab.size == 0:
            ab = np.zeros(3)
        self._bdry_keep = ab
        _check(lib().ibamr_plan_set_physical_boundary(self._h, arr, len(boxes),
                                                      ab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def set_physical_boundary_coefs(self, boxes, coefs):
        """The same boxes with coefficients varying over each codim-1 box: coefs
        lists, for each codim-1 box in order and each array j, the (a, b, g)
        arrays over the box's columns (lower tangential axis fastest; side data:
        one more column along j when j is tangential to the box), each a
        flat sequence or an array of shape (3, n)."""
        import numpy as np
        arr = self._bdry_boxes(boxes)
        flat = [np.ascontiguousarray(c, dtype=np.float64).reshape(-1) for c in coefs]
        cf = np.concatenate(flat) if flat else np.zeros(0)
        keep = cf if cf.size else np.zeros(1)
        self._bdry_keep = keep
        _check(lib().ibamr_plan_set_physical_boundary_coefs(self._h, arr, len(boxes),
                                                            keep.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                            int(cf.size)))

    def accumulate_physical_boundary(self, slot, stream=None):
        """accumulateFromPhysicalBoundaryData on every patch (after a spread)."""
        _check(lib().ibamr_plan_accumulate_physical_boundary(self._h, slot, _stream_ptr(stream)))

    def fill_physical_boundary(self, slot, stream=None):
        """setPhysicalBoundaryConditions on every patch (before an interpolation)."""
        _check(lib().ibamr_plan_fill_physical_boundary(self._h, slot, _stream_ptr(stream)))

    def fill_ghosts(self, slot, stream=None):
        """Same-level ghost fill between the level's patches (periodic wrap included)."""
        _check(lib().ibamr_plan_fill_ghosts(self._h, slot, _stream_ptr(stream)))

    def set_timing(self, on=True):
        """HIP events around each spread / interpolation kernel launch (measurement only)."""
        _check(lib().ibamr_plan_set_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        """Mean kernel durations (ms) and launch counts since the last call."""
        sm, im = ctypes.c_double(), ctypes.c_double()
        ns, ni = ctypes.c_int(), ctypes.c_int()
        _check(lib().ibamr_plan_kernel_times(self._h, ctypes.byref(sm), ctypes.byref(ns), ctypes.byref(im),
                                             ctypes.byref(ni)))
        return {"spread_ms": sm.value, "nspread": ns.value, "interp_ms": im.value, "ninterp": ni.value}


def physical_boundary_boxes(level, patch_index, gcw):
    """Physical boundary boxes of one patch for the Robin boundary operators,
    in SAMRAI's order (codim 1 faces, codim 2 edges, codim 3 corners, each by
    location index), for patches whose physical boundary pieces all lie on
    the level's non-periodic domain sides: a face of the patch on a
    non-periodic domain side gives a codim-1 box (the cell layer outside the
    face); an edge (corner) whose faces are all such faces gives a codim-2
    (codim-3) box, returned as its fill box (gcw cells outward in every normal
    direction, the patch extent along the edge).  Returns [(codim, location,
    lower, upper)].  A caller with SAMRAI passes PatchGeometry's boundary boxes
    instead (getCodimensionBoundaries, getBoundaryFillBox)."""
    p = level.patches[patch_index]
    nd = level.ndim
    lo, hi = list(p.ilower), list(p.iupper)
    phys = [[level.periodic_shift[d] == 0 and lo[d] == level.domain_lower[d],
             level.periodic_shift[d] == 0 and hi[d] == level.domain_upper[d]] for d in range(nd)]
    out = []
    for loc in range(2 * nd):
        d, up = loc // 2, loc % 2
        if phys[d][up]:
            blo, bhi = lo[:], hi[:]
            blo[d] = bhi[d] = (hi[d] + 1) if up else (lo[d] - 1)
            out.append((1, loc, blo, bhi))

    def ghost(d, up):
        return (hi[d] + 1, hi[d] + gcw) if up else (lo[d] - gcw, lo[d] - 1)
    if nd == 3:
        edges = [(0, 1, 2)] * 4 + [(1, 0, 2)] * 4 + [(2, 0, 1)] * 4
        for loc in range(12):
            e, n1, n2 = edges[loc]
            if loc < 4:
                u1, u2 = loc & 1, (loc >> 1) & 1
            elif loc < 8:
                u1, u2 = (loc - 4) >> 1, (loc - 4) & 1
            else:
                u1, u2 = (loc - 8) & 1, (loc - 8) >> 1
            if phys[n1][u1] and phys[n2][u2]:
                blo, bhi = lo[:], hi[:]
                blo[n1], bhi[n1] = ghost(n1, u1)
                blo[n2], bhi[n2] = ghost(n2, u2)
                out.append((2, loc, blo, bhi))
    for loc in range(1 << nd):
        ups = [(loc >> d) & 1 for d in range(nd)]
        if all(phys[d][ups[d]] for d in range(nd)):
            blo, bhi = lo[:], hi[:]
            for d in range(nd):
                blo[d], bhi[d] = ghost(d, ups[d])
            out.append((nd, loc, blo, bhi))
    return out


def permute_rows(dst, src, order, stream=None):
    """dst[k] = src[order[k]] for rows of device tensors (4- or 8-byte elements),
    the VecScatter of LData to the redistributed order."""
    n = order.shape[0]
    depth = (src.numel() // n) if n else 1
    _check(lib().ibamr_permute_rows(_ptr(dst), _ptr(src), _ptr(order), n, depth, src.element_size(),
                                    _stream_ptr(stream)))


def fe_quadrature_points(elem_nodes, phi, JxW, X_nodes, F_nodes=None, f_dtype=None, stream=None):
    """Quadrature points and F*JxW of every element (FEDataManager::spread's
    point source).  elem_nodes [nelem][nbasis] int32, phi [nbasis][nqp],
    JxW [nelem][nqp] or None, X_nodes [nnodes][ndim], F_nodes [nnodes][nvars]
    or None (device tensors, float64).  Returns (X_qp [nelem*nqp][ndim],
    F_qp [nelem*nqp][nvars] or None)."""
    import torch
    nelem, nbasis = elem_nodes.shape
    nqp = phi.shape[1]
    ndim = X_nodes.shape[1]
    if tuple(phi.shape) != (nbasis, nqp) or (JxW is not None and tuple(JxW.shape) != (nelem, nqp)) or \
            (F_nodes is not None and F_nodes.shape[0] != X_nodes.shape[0]):
        raise ValueError("fe_quadrature_points: phi [nbasis][nqp], JxW [nelem][nqp], F_nodes rows = X_nodes rows")
    for t in (elem_nodes, phi, JxW, X_nodes, F_nodes):
        if t is not None and not t.is_contiguous():
            raise ValueError("fe_quadrature_points: contiguous tensors only")
    dev = X_nodes.device
    Xq = torch.empty((max(nelem * nqp, 1), ndim), dtype=torch.float64, device=dev)
    Fq = None
    nvars = 0
    if F_nodes is not None:
        nvars = F_nodes.shape[1]
        Fq = torch.empty((max(nelem * nqp, 1), nvars), dtype=f_dtype or torch.float64, device=dev)
    prec = FP32C if Fq is not None and Fq.dtype == torch.float32 else FP64
    _check(lib().ibamr_fe_quadrature_points(ndim, nvars, nelem, nbasis, nqp, X_nodes.shape[0], _ptr(elem_nodes),
                                            _ptr(phi), _ptr(JxW),
                                            _ptr(X_nodes), _ptr(F_nodes), _ptr(Xq), _ptr(Fq), prec,
                                            _stream_ptr(stream)))
    return Xq[:nelem * nqp], (Fq[:nelem * nqp] if Fq is not None else None)


FE_ELEM_TYPES = ["EDGE2", "EDGE3", "TRI3", "TRI6", "QUAD4", "QUAD8", "QUAD9", "TET4", "TET10", "HEX8", "HEX20",
                 "HEX27"]
FE_QGAUSS, FE_QGRID = 0, 1


def fe_adaptive_quadrature_orders(elem_type, elem_nodes, X_nodes, quad_type, point_density, dx_min, stream=None):
    """getQuadratureKey's adaptive order of every element (device int32 [nelem]):
    elem_type a name of FE_ELEM_TYPES, elem_nodes [nelem][nbasis] int32 and
    X_nodes [nnodes][ndim] float64 device tensors."""
    import torch
    nelem, nbasis = elem_nodes.shape
    out = torch.empty(max(nelem, 1), dtype=torch.int32, device=X_nodes.device)
    _check(lib().ibamr_fe_adaptive_quadrature_orders(FE_ELEM_TYPES.index(elem_type), X_nodes.shape[1], nelem, nbasis,
                                                     X_nodes.shape[0], _ptr(elem_nodes), _ptr(X_nodes), quad_type,
                                                     float(point_density), float(dx_min), _ptr(out),
                                                     _stream_ptr(stream)))
    return out[:nelem]


class FERules:
    """Quadrature rules per element on the device: phis[r] the shape function
    table [nbasis][nqp_r] of rule r, elem_rule[e] the rule of element e.  Holds
    rule_nqp, rule_phi_off, the concatenated tables and qp_off (element e's
    points start at qp_off[e]; nqp_total points)."""

    def __init__(self, phis, elem_rule, device="cuda"):
        import numpy as np
        import torch
        er = np.ascontiguousarray(elem_rule, dtype=np.int32)
        nqp = np.array([np.shape(p)[1] for p in phis], dtype=np.int32)
        off = np.zeros(len(phis), dtype=np.int64)
        off[1:] = np.cumsum([np.size(p) for p in phis])[:-1]
        qp_off = np.zeros(len(er) + 1, dtype=np.int64)
        qp_off[1:] = np.cumsum(nqp[er])
        self.nrules = len(phis)
        self.nqp_total = int(qp_off[-1])
        self.qp_off_host = qp_off
        self.elem_rule = torch.from_numpy(er).to(device)
        self.rule_nqp = torch.from_numpy(nqp).to(device)
        self.rule_phi_off = torch.from_numpy(off).to(device)
        self.phi = torch.from_numpy(np.concatenate([np.asarray(p, dtype=np.float64).reshape(-1) for p in phis])).to(device)
        self.qp_off = torch.from_numpy(qp_off).to(device)


def fe_quadrature_points_rules(elem_nodes, rules, JxW, X_nodes, F_nodes=None, f_dtype=None, stream=None):
    """fe_quadrature_points with a rule per element (FERules); JxW flat
    [nqp_total] (element e's from qp_off[e]) or None.  Returns (X_qp, F_qp)."""
    import torch
    nelem, nbasis = elem_nodes.shape
    ndim = X_nodes.shape[1]
    tot = rules.nqp_total
    if rules.elem_rule.numel() != nelem or (JxW is not None and JxW.numel() != tot):
        raise ValueError("fe_quadrature_points_rules: one rule per element, JxW [nqp_total]")
    dev = X_nodes.device
    Xq = torch.empty((max(tot, 1), ndim), dtype=torch.float64, device=dev)
    Fq = None
    nvars = 0
    if F_nodes is not None:
        nvars = F_nodes.shape[1]
        Fq = torch.empty((max(tot, 1), nvars), dtype=f_dtype or torch.float64, device=dev)
    prec = FP32C if Fq is not None and Fq.dtype == torch.float32 else FP64
    _check(lib().ibamr_fe_quadrature_points_rules(ndim, nvars, nelem, nbasis, X_nodes.shape[0], _ptr(elem_nodes),
                                                  rules.nrules, _ptr(rules.elem_rule), _ptr(rules.rule_nqp),
                                                  _ptr(rules.rule_phi_off), _ptr(rules.phi), _ptr(rules.qp_off), tot,
                                                  _ptr(JxW), _ptr(X_nodes), _ptr(F_nodes), _ptr(Xq), _ptr(Fq), prec,
                                                  _stream_ptr(stream)))
    return Xq[:tot], (Fq[:tot] if Fq is not None else None)


def fe_nodal_weight(V, M_diag, out=None, accumulate=False, stream=None):
    """out = (or +=) V * M_diag[:, None] (nodal quadrature's F (x) dX and U * dX);
    V and out float64 or float32 [nnodes][nvars] device tensors."""
    import torch
    n, nvars = V.shape
    if out is None:
        out = torch.empty_like(V)
    _check(lib().ibamr_fe_nodal_weight(n, nvars, _ptr(V), FP32C if V.dtype == torch.float32 else FP64, _ptr(M_diag),
                                       _ptr(out), FP32C if out.dtype == torch.float32 else FP64,
                                       ACCUMULATE if accumulate else OVERWRITE, _stream_ptr(stream)))
    return out


class FEMesh:
    """Element connectivity on the GPU for FEDataManager::interpWeighted's
    right-hand side and nodal quadrature's diagonal mass (ibamr_fe_mesh_*).
    elem_nodes: host int array [nelem][nbasis]."""

    def __init__(self, elem_nodes, nnodes):
        import numpy as np
        en = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        self.nelem, self.nbasis = en.shape
        self.nnodes = int(nnodes)
        h = ctypes.c_void_p()
        _check(lib().ibamr_fe_mesh_create(ctypes.byref(h), self.nelem, self.nbasis, self.nnodes,
                                          en.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        self._h = h

    def integrate_rhs(self, F_nodes, U_qp, phi, JxW, accumulate=False, stream=None):
        """F_nodes [nnodes][nvars] float64 (device) =/+= sum over elements of
        phi * JxW * U_qp (U_qp [nelem*nqp][nvars], float64 or float32)."""
        import torch
        nqp = phi.shape[1]
        nvars = F_nodes.shape[1]
        if tuple(phi.shape) != (self.nbasis, nqp) or tuple(JxW.shape) != (self.nelem, nqp) or \
                tuple(U_qp.shape) != (self.nelem * nqp, nvars) or F_nodes.shape[0] != self.nnodes or \
                F_nodes.dtype != torch.float64:
            raise ValueError("integrate_rhs: shapes phi [nbasis][nqp], JxW [nelem][nqp], U_qp [nelem*nqp][nvars], "
                             "F_nodes float64 [nnodes][nvars]")
        prec = FP32C if U_qp.dtype == torch.float32 else FP64
        _check(lib().ibamr_fe_integrate_rhs(self._h, nvars, nqp, _ptr(phi), _ptr(JxW), _ptr(U_qp), prec,
                                            _ptr(F_nodes), ACCUMULATE if accumulate else OVERWRITE,
                                            _stream_ptr(stream)))

    def integrate_rhs_rules(self, F_nodes, U_qp, rules, JxW, accumulate=False, stream=None):
        """integrate_rhs with a rule per element (FERules); JxW flat [nqp_total]."""
        import torch
        nvars = F_nodes.shape[1]
        if tuple(U_qp.shape) != (rules.nqp_total, nvars) or JxW.numel() != rules.nqp_total or \
                rules.elem_rule.numel() != self.nelem or F_nodes.dtype != torch.float64:
            raise ValueError("integrate_rhs_rules: U_qp [nqp_total][nvars], JxW [nqp_total], F_nodes float64")
        _check(lib().ibamr_fe_integrate_rhs_rules(self._h, nvars, rules.nrules, _ptr(rules.elem_rule),
                                                  _ptr(rules.rule_nqp), _ptr(rules.rule_phi_off), _ptr(rules.phi),
                                                  _ptr(rules.qp_off), _ptr(JxW), _ptr(U_qp),
                                                  FP32C if U_qp.dtype == torch.float32 else FP64, _ptr(F_nodes),
                                                  ACCUMULATE if accumulate else OVERWRITE, _stream_ptr(stream)))

    def nodal_mass(self, phi, JxW, stream=None):
        """Diagonal mass of a rule (the nodal rule for nodal quadrature): float64 [nnodes]."""
        import torch
        nqp = phi.shape[1]
        if tuple(phi.shape) != (self.nbasis, nqp) or tuple(JxW.shape) != (self.nelem, nqp):
            raise ValueError("nodal_mass: phi [nbasis][nqp], JxW [nelem][nqp]")
        M = torch.empty(max(self.nnodes, 1), dtype=torch.float64, device=phi.device)
        _check(lib().ibamr_fe_nodal_mass(self._h, nqp, _ptr(phi), _ptr(JxW), _ptr(M), _stream_ptr(stream)))
        return M[:self.nnodes]

    def close(self):
        if getattr(self, "_h", None):
            lib().ibamr_fe_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forward_euler_step(X_new, X, U, dt, stream=None):
    """IBMethod::forwardEulerStep: X_new = X + dt U (device tensors; U float64 or float32)."""
    import torch
    prec = FP64 if U.dtype == torch.float64 else FP32C
    assert X.dtype == torch.float64 and X_new.dtype == torch.float64 and X.numel() == U.numel() == X_new.numel()
    _check(lib().ibamr_forward_euler_step(_ptr(X_new), _ptr(X), _ptr(U), prec, float(dt), X.numel(),
                                          _stream_ptr(stream)))


def forward_euler_step_periodic(X_new, X, U, dt, level, stream=None):
    """X_new = X + dt U, then positions moved back into the level's domain by one
    period along its periodic axes (device tensors [n][ndim])."""
    import torch
    prec = FP64 if U.dtype == torch.float64 else FP32C
    nd = level.ndim
    assert X.dtype == torch.float64 and X_new.dtype == torch.float64 and X.numel() == U.numel() == X_new.numel()
    lo = (ctypes.c_double * nd)(*level.domain_x_lower)
    hi = (ctypes.c_double * nd)(*level.domain_x_upper)
    per = (ctypes.c_int * nd)(*[1 if s else 0 for s in level.periodic_shift])
    _check(lib().ibamr_forward_euler_step_periodic(_ptr(X_new), _ptr(X), _ptr(U), prec, float(dt), X.numel() // nd, nd,
                                                   lo, hi, per, _stream_ptr(stream)))


# --------------------------------------------------------------------------- S0: Fortran-kernel drop-in

def _np_d(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _np_i(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def s0_interp(kernel, dx, x_lower, ilower, iupper, nugc, u, X, V, indices, Xshift, depth=1):
    """lagrangian_<kernel>_interp{2,3}d on host arrays (runs on the GPU); V is
    updated in place (float64, C-contiguous, shape (N, depth))."""
    import numpy as np
    ndim = len(ilower)
    assert V.dtype == np.float64 and V.flags.c_contiguous
    keep = [_np_d(dx), _np_d(x_lower), _np_i(ilower), _np_i(iupper), _np_i(nugc), _np_d(u), _np_i(indices),
            _np_d(Xshift), _np_d(X)]
    _check(lib().ibamr_s0_interp(kernel.encode(), ndim, keep[0][1], keep[1][1], depth, keep[2][1], keep[3][1],
                                 keep[4][1], keep[5][1], keep[6][1], keep[7][1], len(keep[6][0]), keep[8][1],
                                 V.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return V


def s0_spread(kernel, dx, x_lower, ilower, iupper, nugc, u, X, V, indices, Xshift, depth=1):
    """lagrangian_<kernel>_spread{2,3}d on host arrays (runs on the GPU); u is
    accumulated in place (float64, C-contiguous)."""
    import numpy as np
    ndim = len(ilower)
    assert u.dtype == np.float64 and u.flags.c_contiguous
    keep = [_np_d(dx), _np_d(x_lower), _np_i(indices), _np_d(Xshift), _np_d(X), _np_d(V), _np_i(ilower),
            _np_i(iupper), _np_i(nugc)]
    _check(lib().ibamr_s0_spread(kernel.encode(), ndim, keep[0][1], keep[1][1], depth, keep[2][1], keep[3][1],
                                 len(keep[2][0]), keep[4][1], keep[5][1], keep[6][1], keep[7][1], keep[8][1],
                                 u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return u
