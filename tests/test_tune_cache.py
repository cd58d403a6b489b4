"""Autotuner cache (cs_tune_*): keys, persistence and lookup -- host only, no GPU."""
import ctypes
import os

import numpy as np
import pytest

from convstencil_amd import _lib, Layout, shape_offsets, dense_weights
from convstencil_amd.api import _options
from oracle import oracle


@pytest.fixture()
def lib(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.txt"
    monkeypatch.setenv("CONVSTENCIL_TUNE_CACHE", str(path))
    lib = _lib.load()
    lib.cs_tune_cache_reset()
    yield lib, path
    monkeypatch.delenv("CONVSTENCIL_TUNE_CACHE")
    lib.cs_tune_cache_reset()


def _lookup(lib, key):
    v, ms = ctypes.c_int32(), ctypes.c_double()
    st = lib.cs_tune_lookup(key, ctypes.byref(v), ctypes.byref(ms))
    return st, v.value, ms.value


def test_store_lookup_persist(lib):
    lib, path = lib
    assert lib.cs_tune_cache_path().decode() == str(path)
    assert _lookup(lib, b"d2_r1_s0_p0_n1x64x64_g0_f0")[0] == _lib.CS_ERR_UNSUPPORTED
    assert lib.cs_tune_store(b"d2_r1_s0_p0_n1x64x64_g0_f0", 111108, 0.25) == _lib.CS_OK
    assert _lookup(lib, b"d2_r1_s0_p0_n1x64x64_g0_f0") == (_lib.CS_OK, 111108, 0.25)
    # the file was created (with its directory) and is read back after a reset
    assert path.exists()
    lib.cs_tune_cache_reset()
    assert _lookup(lib, b"d2_r1_s0_p0_n1x64x64_g0_f0") == (_lib.CS_OK, 111108, 0.25)
    # a later entry for the same key wins, in memory and in the file
    assert lib.cs_tune_store(b"d2_r1_s0_p0_n1x64x64_g0_f0", 11104, 0.5) == _lib.CS_OK
    lib.cs_tune_cache_reset()
    assert _lookup(lib, b"d2_r1_s0_p0_n1x64x64_g0_f0") == (_lib.CS_OK, 11104, 0.5)
    assert len(path.read_text().splitlines()) == 2


def test_invalid_keys(lib):
    lib, path = lib
    assert lib.cs_tune_store(b"has space", 1, 1.0) == _lib.CS_ERR_INVALID
    assert lib.cs_tune_store(b"", 1, 1.0) == _lib.CS_ERR_INVALID
    assert lib.cs_tune_store(b"ok_key", 0, 1.0) == _lib.CS_ERR_INVALID
    assert _lookup(lib, b"bad key")[0] == _lib.CS_ERR_INVALID
    assert not path.exists()


def test_memory_only(monkeypatch, tmp_path):
    monkeypatch.setenv("CONVSTENCIL_TUNE_CACHE", "")
    lib = _lib.load()
    lib.cs_tune_cache_reset()
    assert lib.cs_tune_cache_path().decode() == ""
    assert lib.cs_tune_store(b"k1", 5, 1.0) == _lib.CS_OK
    assert _lookup(lib, b"k1")[:2] == (_lib.CS_OK, 5)
    lib.cs_tune_cache_reset()
    assert _lookup(lib, b"k1")[0] == _lib.CS_ERR_UNSUPPORTED
    monkeypatch.delenv("CONVSTENCIL_TUNE_CACHE")
    lib.cs_tune_cache_reset()


@pytest.mark.parametrize("name,sizes,prec,graph,fusion,key", [
    ("box2d1r", [10240, 10240], "fp64", False, 0, "d2_r1_s0_p0_n1x10240x10240_g0_f0"),
    ("star2d3r", [8192, 8192], "bf16x3", True, 0, "d2_r3_s1_p1_n1x8192x8192_g1_f0"),
    ("box2d3r", [300, 200], "fp16x2", False, 2, "d2_r3_s0_p3_n1x300x200_g0_f2"),
    ("star3d1r", [64, 32, 16], "fp64", True, 0, "d3_r1_s1_p0_n64x32x16_g1_f0"),
    ("1d1r", [1 << 20], "fp64", False, 0, "d1_r1_s0_p0_n1x1x1048576_g0_f0"),
])
def test_key(name, sizes, prec, graph, fusion, key):
    from convstencil_amd import SHAPES
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo(sizes, r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 3)
    dense = np.ascontiguousarray(dense_weights(ndim, r, offs, w))
    st = _lib.cs_stencil(ndim, r, dense.ctypes.data_as(_lib.c_double_p))
    o = _options(precision=prec, use_graph=graph, fusion_depth=fusion)
    buf = ctypes.create_string_buffer(160)
    lib = _lib.load()
    assert lib.cs_tune_key(ctypes.byref(L.to_c()), ctypes.byref(st), ctypes.byref(o), buf, 160) == _lib.CS_OK
    assert buf.value.decode() == key
