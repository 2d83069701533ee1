"""The C-ABI library loads on a machine without a GPU and exports every
function include/*.h declares; argument checking that needs no device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("shcfd.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(shcfd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def _lib():
    from shared_cfd_amd import native

    return native.lib()


def test_library_exports_every_declared_symbol():
    L = _lib()
    declared = _declared()
    assert len(declared) >= 30
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_lists_every_export():
    from shared_cfd_amd import native

    assert sorted(native.EXPORTS) == _declared()


def test_strerror_and_argument_errors_without_gpu():
    L = _lib()
    assert L.shcfd_strerror(0) == b"ok"
    assert L.shcfd_strerror(5).startswith(b"Omega")
    out = ctypes.c_void_p()
    # invalid sizes are rejected before any device call
    ni = (ctypes.c_int * 1)(1)
    nj = (ctypes.c_int * 1)(4)
    assert L.shcfd_create(1, ni, nj, 1, None, ctypes.byref(out)) != 0
    assert L.shcfd_create(0, ni, nj, 1, None, ctypes.byref(out)) != 0
    assert L.shcfd_destroy(None) != 0
    assert L.shcfd_set_w(None) != 0
    assert L.shcfd_version() >= 100


def test_oracle_is_not_linked_into_the_product():
    """The product library must not depend on the CPU oracle."""
    so = os.path.join(ROOT, "shared_cfd_amd", "libshcfd.so")
    data = open(so, "rb").read()
    assert b"or_create" not in data and b"liboracle" not in data
    src = os.path.join(ROOT, "shared_cfd_amd")
    for dp, _, files in os.walk(src):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".f90")):
                txt = open(os.path.join(dp, f), errors="ignore").read()
                assert "oracle" not in txt.replace("oracle/", ""), os.path.join(dp, f)
