"""The C ABI library loads and exports every function include/fastice_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fi_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("fi_model_create", "fi_advance_iterations", "fi_compute_residuals", "fi_solve",
                 "fi_omega_from_psi", "fi_levelset_from_dem"):
        assert must in names


def test_library_exports_every_declared_symbol(pkg):
    lib = C.CDLL(pkg._abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_descriptor_layout(pkg):
    lib = pkg._abi.lib()
    assert lib.fi_abi_version() == pkg._abi.FI_ABI_VERSION
    # fi_model_desc layout: natural alignment, doubles 8-aligned
    d = pkg._abi.fi_model_desc
    assert d.n.offset == 4 and d.origin.offset == 16
    assert C.sizeof(pkg._abi.fi_rheology) == 56


def test_bad_descriptor_is_rejected_without_gpu(pkg):
    import numpy as np
    A = pkg._abi
    d = A.fi_model_desc()
    d.ndims = 4
    h = C.c_void_p()
    rc = A.lib().fi_model_create(C.byref(d), C.byref(h))
    assert rc == A.FI_ERR_ARG
    assert b"ndims" in A.lib().fi_last_error()
    assert not h.value
    assert A.lib().fi_advance_iterations(None, 1) == A.FI_ERR_ARG
    out = np.zeros(4)
    assert A.lib().fi_compute_residuals(None, A.dptr(out)) == A.FI_ERR_ARG
