"""Lagrangian structures: the IBStandardInitializer input files and the GPU
force generator.

read_structure(base) reads base.vertex / .spring / .beam / .target in the
formats of src/IB/IBStandardInitializer.cpp (readVertexFiles :184-294,
readSpringFiles :297-528, readBeamFiles :766-1002, readTargetPointFiles):
a count on the first line, then one entry per line, text after '!', '#' or
'%' ignored.  Vertex lines: NDIM coordinates.  Spring lines: node, node,
stiffness, rest length[, force function index[, further parameters]] (the
parameters of a spring are stiffness, rest length and the further values,
IBStandardInitializer.cpp:397-434).
Beam lines: previous, current, next node, bending rigidity[, NDIM curvature
components].  Target lines: node, stiffness[, damping].

Forces wraps ibamr_forces_* (IBStandardForceGen springs, beams, target points,
the spring force function registry).
"""
import ctypes
import os
import re

import numpy as np

import ibamr_amd as ia

SPRING_HOOKEAN, SPRING_LINEAR_STRAIN, SPRING_POLYNOMIAL = 0, 1, 2


def _lines(path):
    with open(path) as f:
        for raw in f:
            line = re.split(r"[!#%]", raw, maxsplit=1)[0].strip()
            if line:
                yield line.split()


def _read(path):
    it = _lines(path)
    n = int(next(it)[0])
    rows = [next(it) for _ in range(n)]
    return rows


def read_structure(base, ndim):
    """Returns a dict of numpy arrays: X (n, ndim); springs (master, slave,
    kappa, rest); spring_functions (force function index, parameter offsets,
    parameters) of every spring; beams (prev, mastr, next, rigidity,
    curvature); targets (node, kappa, eta, X0).  Missing files give None."""
    X = np.array([[float(v) for v in r[:ndim]] for r in _read(base + ".vertex")], dtype=np.float64)
    out = {"X": X, "springs": None, "spring_functions": None, "beams": None, "targets": None}
    if os.path.exists(base + ".spring"):
        rows = _read(base + ".spring")
        idx = np.array([int(r[4]) if len(r) > 4 else 0 for r in rows], dtype=np.int32)
        par = [[float(v) for v in r[2:4]] + [float(v) for v in r[5:]] for r in rows]
        off = np.zeros(len(rows) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(p) for p in par])
        out["spring_functions"] = (idx, off, np.array([v for p in par for v in p], dtype=np.float64))
        out["springs"] = (np.array([int(r[0]) for r in rows], dtype=np.int32),
                          np.array([int(r[1]) for r in rows], dtype=np.int32),
                          np.array([float(r[2]) for r in rows]), np.array([float(r[3]) for r in rows]))
    if os.path.exists(base + ".beam"):
        rows = _read(base + ".beam")
        curv = np.array([[float(v) for v in r[4:4 + ndim]] if len(r) >= 4 + ndim else [0.0] * ndim for r in rows])
        out["beams"] = (np.array([int(r[0]) for r in rows], dtype=np.int32),
                        np.array([int(r[1]) for r in rows], dtype=np.int32),
                        np.array([int(r[2]) for r in rows], dtype=np.int32),
                        np.array([float(r[3]) for r in rows]), curv.reshape(-1, ndim))
    if os.path.exists(base + ".target"):
        rows = _read(base + ".target")
        node = np.array([int(r[0]) for r in rows], dtype=np.int32)
        out["targets"] = (node, np.array([float(r[1]) for r in rows]),
                          np.array([float(r[2]) if len(r) > 2 else 0.0 for r in rows]), X[node].copy())
    return out


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if a is not None and len(a) else None


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None and len(a) else None


class Forces:
    """Spring, beam and target-point forces of one set of n nodes on the GPU.

    springs = (master, slave, kappa, rest); beams = (prev, master, next,
    rigidity, curvature[nb][ndim]); targets = (node, kappa, eta, X0[nt][ndim]).
    Springs start as the Hookean default (force function 0); registry maps
    force function indices to laws (SPRING_HOOKEAN, SPRING_LINEAR_STRAIN,
    SPRING_POLYNOMIAL) and spring_functions = (index per spring, parameter
    offsets, parameters) then sets each spring's law and parameters."""

    def __init__(self, ndim, npoints, springs=None, beams=None, targets=None, registry=None, spring_functions=None):
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        s = [i32(springs[0]), i32(springs[1]), f64(springs[2]), f64(springs[3])] if springs is not None else [None] * 4
        if beams is not None:
            b = [i32(beams[1]), i32(beams[2]), i32(beams[0]), f64(beams[3]), f64(np.reshape(beams[4], (-1, ndim)))]
        else:
            b = [None] * 5
        t = [i32(targets[0]), f64(targets[1]), f64(targets[2]), f64(np.reshape(targets[3], (-1, ndim)))] \
            if targets is not None else [None] * 4
        self._keep = (s, b, t)
        ns = len(s[0]) if s[0] is not None else 0
        nb = len(b[0]) if b[0] is not None else 0
        nt = len(t[0]) if t[0] is not None else 0
        h = ctypes.c_void_p()
        ia._check(ia.lib().ibamr_forces_create(ctypes.byref(h), ndim, npoints, ns, _ip(s[0]), _ip(s[1]), _dp(s[2]),
                                               _dp(s[3]), nb, _ip(b[0]), _ip(b[1]), _ip(b[2]), _dp(b[3]), _dp(b[4]),
                                               nt, _ip(t[0]), _dp(t[1]), _dp(t[2]), _dp(t[3])))
        self._h = h
        for idx, law in (registry or {}).items():
            self.register_spring_function(idx, law)
        if spring_functions is not None:
            self.set_spring_parameters(*spring_functions)

    def register_spring_function(self, force_fcn_index, law):
        """IBStandardForceGen::registerSpringForceFunction with a built-in law."""
        ia._check(ia.lib().ibamr_forces_register_spring_function(self._h, int(force_fcn_index), int(law)))

    def set_spring_parameters(self, force_fcn_idx, offsets, params):
        """Each spring's force function index and parameters
        params[offsets[k]:offsets[k+1]], resolved through the registry now."""
        i = np.ascontiguousarray(force_fcn_idx, dtype=np.int32)
        o = np.ascontiguousarray(offsets, dtype=np.int32)
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._spar = (i, o, p)
        ia._check(ia.lib().ibamr_forces_set_spring_parameters(self._h, _ip(i), _ip(o), _dp(p)))

    def compute(self, F, X, U=None, accumulate=False, stream=None):
        """F (device tensor, float64 or float32) += / = the force at X (float64);
        U (float64 or float32, or None) damps the target points."""
        import torch
        prec = ia.FP64 if F.dtype == torch.float64 else ia.FP32C
        uprec = ia.FP32C if U is not None and U.dtype == torch.float32 else ia.FP64
        mode = ia.ACCUMULATE if accumulate else ia.OVERWRITE
        ia._check(ia.lib().ibamr_forces_compute(self._h, ia._ptr(F), prec, ia._ptr(X), ia._ptr(U), uprec, mode,
                                                ia._stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None):
            ia.lib().ibamr_forces_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
