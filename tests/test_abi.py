"""The C ABI: libconvstencil.so loads (no GPU needed) and exports every function that
include/convstencil.h and include/convstencil_debug.h declare, and the ctypes binding
(convstencil_amd/_lib.py) binds every function of the public header."""
import os
import re

import pytest

from convstencil_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # comments
    text = re.sub(r"#.*", "", text)  # preprocessor lines
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
        name = m.group(1)
        # a declaration's name follows a return type (not a keyword or a call inside a macro)
        before = text[:m.start()].rstrip()
        if before.endswith(("*", "int", "void", "cs_status", "int32_t", "int64_t", "double", "char")):
            names.add(name)
    return sorted(names)


PUBLIC = _declared("convstencil.h")
DEBUG = _declared("convstencil_debug.h")


def test_headers_parse():
    assert "cs_plan_create" in PUBLIC and "gpu_box_2d1r" in PUBLIC and "cs_run_2d_noalloc" in PUBLIC
    assert "cs_debug_variant_info" in DEBUG
    assert len(PUBLIC) >= 40


@pytest.mark.parametrize("name", PUBLIC + DEBUG)
def test_symbol_exported(name):
    lib = _lib.load()
    assert hasattr(lib, name), name


@pytest.mark.parametrize("name", PUBLIC)
def test_public_symbol_bound(name):
    assert name in _lib.SIGNATURES, name
