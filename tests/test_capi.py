"""CPU checks of the C ABI: the built library loads and exports every symbol
the public headers declare, and the registry answers like
LEInteractor::isKnownKernel/getStencilSize/getMinimumGhostWidth
(LEInteractor.cpp:2035-2114).  No GPU computation is called."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    import ibamr_amd
    if not os.path.exists(ibamr_amd.LIB_PATH):
        from ibamr_amd import build
        build.build()
    return ibamr_amd.lib()


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", txt, flags=re.M):
                name = m.group(1)
                if name not in ("if", "for", "while", "return", "sizeof"):
                    names.add(name)
    return names


def test_header_symbols_exported(L):
    import ibamr_amd
    names = _declared()
    assert "ibamr_plan_spread" in names and "ibamr_plan_interp" in names
    out = subprocess.run(["nm", "-D", "--defined-only", ibamr_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing
    for n in ibamr_amd.C_API:
        assert n in names, n
        getattr(L, n)


def test_registry(L):
    import ibamr_amd as ia
    for k, s, g in [("IB_4", 4, 3), ("IB_6", 6, 4), ("BSPLINE_3", 4, 3), ("PIECEWISE_LINEAR", 2, 2)]:
        assert ia.is_known_kernel(k)
        assert ia.get_stencil_size(k) == s
        assert ia.get_minimum_ghost_width(k) == g
    assert not ia.is_known_kernel("IB_7")
    with pytest.raises(ia.IBAMRError) as e:
        ia.get_stencil_size("IB_7")
    assert e.value.code == ia.IBAMR_ERR_BAD_KERNEL


def _level3(n=8, periodic=True):
    import ibamr_amd as ia
    lv = ia.Level.uniform_periodic(n, 3)
    if not periodic:
        lv.periodic_shift = [0, 0, 0]
    return lv


@pytest.mark.parametrize("kw,code_name", [
    (dict(kernel="IB_7"), "IBAMR_ERR_BAD_KERNEL"),
    (dict(kernel="PIECEWISE_LINEAR"), "IBAMR_ERR_BAD_KERNEL"),
    (dict(kernel="IB_6", data_gcw=[3, 3, 3]), "IBAMR_ERR_GHOST_WIDTH"),
    (dict(kernel="IB_4", data_gcw=[3, 2, 3]), "IBAMR_ERR_GHOST_WIDTH"),
    (dict(kernel="IB_4", idx_gcw=2), "IBAMR_ERR_GHOST_WIDTH"),
    (dict(kernel="IB_4", depth=2), "IBAMR_ERR_DEPTH"),
])
def test_plan_create_rejects(L, kw, code_name):
    """Host-side validation (the reference's TBOX_ERROR cases, LEInteractor.cpp:
    2014, 4491-4498, 2197-2201) returns a status before touching the GPU."""
    import ibamr_amd as ia
    with pytest.raises(ia.IBAMRError) as e:
        ia.LevelPlan(_level3(), centering=ia.SIDE, max_points=10, **kw)
    assert e.value.code == getattr(ia, code_name)


def test_plan_create_rejects_small_periodic_domain(L):
    import ibamr_amd as ia
    with pytest.raises(ia.IBAMRError) as e:
        ia.LevelPlan(_level3(n=3), kernel="IB_6", centering=ia.CELL, max_points=10)
    assert e.value.code == ia.IBAMR_ERR_INVALID
