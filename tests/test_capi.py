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
    (dict(kernel="IB_4_W8"), "IBAMR_ERR_BAD_KERNEL"),
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


def test_registry_full_reference_table(L):
    """All 22 names of LEInteractor::isKnownKernel with the stencil sizes of
    getStencilSize (LEInteractor.cpp:2035-2108); names without a GPU kernel are
    known but refused by plan creation."""
    import ibamr_amd as ia
    table = {"BSPLINE_3": 4, "BSPLINE_4": 4, "BSPLINE_5": 6, "BSPLINE_6": 6, "COMPOSITE_BSPLINE_23": 4,
             "COMPOSITE_BSPLINE_32": 4, "COMPOSITE_BSPLINE_34": 4, "COMPOSITE_BSPLINE_43": 4,
             "COMPOSITE_BSPLINE_45": 5, "COMPOSITE_BSPLINE_54": 5, "COMPOSITE_BSPLINE_56": 6,
             "COMPOSITE_BSPLINE_65": 6, "DISCONTINUOUS_LINEAR": 2, "PIECEWISE_CONSTANT": 1, "PIECEWISE_LINEAR": 2,
             "PIECEWISE_CUBIC": 4, "IB_3": 4, "IB_4": 4, "IB_4_W8": 8, "IB_5": 6, "IB_6": 6}
    for k, s in table.items():
        assert ia.is_known_kernel(k), k
        assert ia.get_stencil_size(k) == s, k
        assert ia.get_minimum_ghost_width(k) == s // 2 + 1, k
    assert ia.is_known_kernel("USER_DEFINED")
    for k in ("IB_3", "COMPOSITE_BSPLINE_43", "PIECEWISE_CUBIC"):
        with pytest.raises(ia.IBAMRError) as e:
            ia.LevelPlan(_level3(), k, centering=ia.SIDE, max_points=10)
        assert e.value.code == ia.IBAMR_ERR_BAD_KERNEL


def test_cpp_driver_reports_reference_errors():
    """The C++ layer raises the reference's TBOX_ERROR text (LEInteractor.cpp:
    2103-2104) before any GPU call; bad usage exits 2."""
    drv = os.path.join(ROOT, "tests", "cpp", "bin", "interpolate_01")
    if not os.path.exists(drv):
        from ibamr_amd import build
        build.build_drivers()
    r = subprocess.run([drv, "interp", "IB_7", os.devnull], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "LEInteractor::getStencilSize()" in r.stderr and "Unknown kernel function IB_7" in r.stderr
    r = subprocess.run([drv], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def test_patch_plan_rejects_bad_box(L):
    import ctypes
    import ibamr_amd as ia
    lv = ia.Level([0.1] * 3, [0] * 3, [9] * 3, [0.0] * 3, [1.0] * 3, [0] * 3,
                  [ia.Patch([0] * 3, [9] * 3, [0.0] * 3, [1.0] * 3)])
    with pytest.raises(ia.IBAMRError) as e:
        ia.LevelPlan(lv, "IB_4", ia.SIDE, max_points=10, box=([5] * 3, [4] * 3))
    assert e.value.code == ia.IBAMR_ERR_INVALID


def test_forces_create_rejects_bad_elements(L):
    """Element validation happens on the host, before any device call."""
    import numpy as np

    import ibamr_amd as ia
    from ibamr_amd import structure

    with pytest.raises(ia.IBAMRError, match="spring 1: bad node index"):
        structure.Forces(2, 3, springs=([0, 2], [1, 3], [1.0, 1.0], [0.0, 0.0]))
    with pytest.raises(ia.IBAMRError, match="spring 0: bad node index"):
        structure.Forces(2, 3, springs=([1], [1], [1.0], [0.0]))
    with pytest.raises(ia.IBAMRError, match="beam 0: bad node index"):
        structure.Forces(3, 3, beams=([0], [0], [1], [1.0], np.zeros((1, 3))))
    with pytest.raises(ia.IBAMRError, match="target point 0: bad node index"):
        structure.Forces(3, 3, targets=([-1], [1.0], [0.0], np.zeros((1, 3))))
    h = ctypes.c_void_p()
    rc = ia.lib().ibamr_forces_create(ctypes.byref(h), 4, 3, 0, None, None, None, None, 0, None, None, None, None,
                                      None, 0, None, None, None, None)
    assert rc != 0 and "ndim" in ia.lib().ibamr_last_error().decode()


def test_new_entry_points_reject_bad_arguments_before_device_work(L):
    """Argument errors of the FE rule / nodal quadrature, spring law and varying
    Robin coefficient entry points come back as IBAMR_ERR_INVALID with a message,
    before any device call (so this runs without a GPU)."""
    import ibamr_amd as ia
    cases = [
        (lambda: L.ibamr_fe_adaptive_quadrature_orders(99, 3, 1, 8, 8, None, None, 0, 1.0, 0.1, None, None),
         "unknown element type"),
        (lambda: L.ibamr_fe_adaptive_quadrature_orders(ia.FE_ELEM_TYPES.index("HEX8"), 3, 1, 8, 8, None, None, 5, 1.0,
                                                       0.1, None, None), "only for QGAUSS or QGRID"),
        (lambda: L.ibamr_fe_adaptive_quadrature_orders(ia.FE_ELEM_TYPES.index("HEX27"), 3, 1, 8, 8, None, None, 0,
                                                       1.0, 0.1, None, None), "below the element type"),
        (lambda: L.ibamr_fe_adaptive_quadrature_orders(ia.FE_ELEM_TYPES.index("TET4"), 3, 1, 4, 4, None, None, 1, 1.0,
                                                       0.0, None, None), "dx_min must be positive"),
        (lambda: L.ibamr_fe_nodal_weight(4, 3, None, 0, None, None, 0, 7, None), "bad mode"),
        (lambda: L.ibamr_fe_nodal_weight(4, 3, None, 3, None, None, 0, 0, None), "bad precision"),
        (lambda: L.ibamr_fe_nodal_mass(None, 8, None, None, None, None), "null mesh"),
        (lambda: L.ibamr_fe_quadrature_points_rules(4, 1, 1, 8, 8, None, 1, None, None, None, None, None, 8, None,
                                                    None, None, None, None, 0, None), "ndim must be 2 or 3"),
        (lambda: L.ibamr_fe_quadrature_points_rules(3, 1, 1, 8, 8, None, 0, None, None, None, None, None, 8, None,
                                                    None, None, None, None, 0, None), "bad element or rule sizes"),
        (lambda: L.ibamr_fe_integrate_rhs_rules(None, 3, 1, None, None, None, None, None, None, None, 0, None, 0,
                                                None), "null mesh"),
        (lambda: L.ibamr_forces_register_spring_function(None, 0, 1), "null force object"),
        (lambda: L.ibamr_forces_set_spring_parameters(None, None, None, None), "null force object"),
        (lambda: L.ibamr_plan_set_physical_boundary_coefs(None, None, 0, None, 0), "null plan"),
    ]
    for call, msg in cases:
        with pytest.raises(ia.IBAMRError, match=msg) as e:
            ia._check(call())
        assert e.value.code == ia.IBAMR_ERR_INVALID, msg
