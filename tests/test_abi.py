"""The C ABI library loads and exports every entry point include/samrai_fac.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(sfac_[a-z0-9_]+)\s*\(", text,
                             flags=re.M):
            names.add(m.group(1))
        # link-compatible Fortran twins (samrai_f77.h): lower case, trailing underscore
        for m in re.finditer(r"^void\s+([a-z0-9]+_)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_api():
    names = declared_symbols()
    for must in ("sfac_solver_solve", "sfac_rbgswithfluxmax3d", "sfac_ewingfixflux2d",
                 "sfac_cartclinrefcelldoub", "sfac_ops_smooth", "sfac_settype3cells3d"):
        assert must in names
    assert len(names) >= 55
    f77 = {n for n in names if n.endswith("_")}
    for must in ("rbgswithfluxmaxvardcvarsf3d_", "compresconsca2d_", "ewingfixfluxcondc3d_",
                 "settype1cells3d_", "compfluxvardc2d_"):
        assert must in f77
    assert len(f77) == 38


def test_library_exports_every_declared_symbol():
    from samrai_amd import fac
    lib = ctypes.CDLL(fac.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_errors_without_throwing():
    from samrai_amd import fac
    L = fac.lib()
    # bad dimension: the C++ side throws, the ABI returns a status
    import numpy as np
    lo = np.zeros(3, np.int32)
    hi = np.ones(3, np.int32)
    x = np.zeros(3)
    h = ctypes.c_void_p()
    st = L.sfac_hierarchy_create(5, lo.ctypes.data_as(fac._ip), hi.ctypes.data_as(fac._ip),
                                 x.ctypes.data_as(fac._dp), x.ctypes.data_as(fac._dp),
                                 ctypes.byref(h))
    assert st != 0
    assert b"dim" in L.sfac_last_error()
