"""Lagrangian structures: the IBStandardInitializer input files and the GPU
force generator.

read_structure(base) reads base.vertex / .spring / .beam / .target in the
formats of src/IB/IBStandardInitializer.cpp (readVertexFiles :184-294,
readSpringFiles :297-528, readBeamFiles :766-1002, readTargetPointFiles):
a count on the first line, then one entry per line, text after '!', '#' or
'%' ignored.  Vertex lines: NDIM coordinates.  Spring lines: node, node,
stiffness, rest length[, force function index (only 0, the Hookean default)].
Beam lines: previous, current, next node, bending rigidity[, NDIM curvature
components].  Target lines: node, stiffness[, damping].

Forces wraps ibamr_forces_* (IBStandardForceGen springs, beams, target points).
"""
import ctypes
import os
import re

import numpy as np

import ibamr_amd as ia


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
    kappa, rest); beams (prev, mastr, next, rigidity, curvature); targets
    (node, kappa, eta, X0).  Missing files give empty arrays."""
    X = np.array([[float(v) for v in r[:ndim]] for r in _read(base + ".vertex")], dtype=np.float64)
    out = {"X": X, "springs": None, "beams": None, "targets": None}
    if os.path.exists(base + ".spring"):
        rows = _read(base + ".spring")
        for r in rows:
            if len(r) > 4 and int(r[4]) != 0:
                raise ValueError("only the default (Hookean) spring force function is supported")
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
    rigidity, curvature[nb][ndim]); targets = (node, kappa, eta, X0[nt][ndim])."""

    def __init__(self, ndim, npoints, springs=None, beams=None, targets=None):
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
