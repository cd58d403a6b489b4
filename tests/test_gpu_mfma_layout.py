"""FP64 MFMA lane layouts on the device (the layouts the kernels and the CPU
emulation in test_kernel_emulation.py assume)."""
import ctypes

import numpy as np
import pytest

from convstencil_amd import _lib

pytestmark = pytest.mark.gpu


def _probe(which):
    lib = _lib.load()
    lib.cs_debug_mfma_layout.restype = ctypes.c_int32
    lib.cs_debug_mfma_layout.argtypes = [ctypes.c_int32, _lib.c_double_p]
    n = 64 * 64 * (4 if which == 0 else 1)
    out = np.zeros(n)
    assert lib.cs_debug_mfma_layout(which, out.ctypes.data_as(_lib.c_double_p)) == 0
    return out


def code(q):
    return 1.0 + q / 64.0


def test_f64_16x16x4_layout(gpu):
    raw = _probe(0).reshape(64, 64, 4)
    want = np.zeros_like(raw)
    for p in range(64):
        i, k = p % 16, p // 16            # A[i][k] at lane i + 16k
        for r in range(64):
            for v in range(4):
                row, col = 4 * v + r // 16, r % 16   # D[4v + l/16][l%16]
                if row == i:
                    want[p, r, v] = code(16 * k + col)  # B[k][j] at lane 16k + j
    assert np.array_equal(raw, want)


def test_f64_4x4x4_layout(gpu):
    raw = _probe(1).reshape(64, 64)
    want = np.zeros_like(raw)
    for p in range(64):
        b, i, k = (p >> 2) & 3, p & 3, p >> 4
        for r in range(64):
            if (r >> 2) & 3 == b and (r >> 4) == i:
                j = r & 3
                want[p, r] = code(16 * k + 4 * b + j)
    assert np.array_equal(raw, want)
