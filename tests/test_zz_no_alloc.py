"""The allocation-free entry point cs_run_2d_noalloc, called through the C ABI on device
buffers the caller owns (allocated here through the HIP runtime), against the plan API (cs_plan_run), which
allocates its own buffers and B fragments: the same kernels, so the results must agree bit
for bit (and match the CPU oracle to 1e-12).  Two calls in a row, the second on a smaller
grid: the library's scratch array keeps the first call's fragments, which the second call
must replace."""
import ctypes

import numpy as np
import pytest

from convstencil_amd import _lib, Plan, Layout, shape_offsets, dense_weights, SHAPES
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime through ctypes: device buffers owned by the test (the caller)."""

    def __init__(self):
        import ctypes.util
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), src, a.nbytes, 2) == 0


def _noalloc(L, R, dense, x, steps, fusion=0):
    hip = _Hip()
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    a, b = hip.alloc(x.nbytes), hip.alloc(x.nbytes)
    try:
        hip.h2d(a, x)
        assert hip.lib.hipMemset(b, 0, x.nbytes) == 0
        w = np.ascontiguousarray(dense, dtype=np.float64)
        st = _lib.cs_stencil(2, R, w.ctypes.data_as(_lib.c_double_p))
        opt = _lib.cs_options()
        opt.precision = _lib.CS_PREC_FP64
        opt.fusion_depth = fusion
        Lc = L.to_c()
        fin = ctypes.c_int32(-1)
        rc = lib.cs_run_2d_noalloc(ctypes.byref(Lc), ctypes.byref(st), a, b, int(steps), ctypes.byref(opt), None,
                                   ctypes.byref(fin))
        assert rc == _lib.CS_OK, rc
        assert hip.lib.hipDeviceSynchronize() == 0
        assert fin.value in (0, 1)
        out = np.empty_like(x)
        hip.d2h(out, a if fin.value == 0 else b)
        return out
    finally:
        hip.lib.hipFree(a)
        hip.lib.hipFree(b)


def _case(name, sizes, steps, seed, fusion=0):
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo(sizes, r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), seed)
    dense = dense_weights(2, r, offs, w)
    x = oracle.fill_data(L, seed)
    got = _noalloc(L, r, dense, x, steps, fusion)
    with Plan(L, r, dense, fusion_depth=fusion) as p:
        p.upload(x)
        p.run(steps)
        ref = p.download()
    assert np.array_equal(got, ref), (name, sizes, steps)
    want = oracle.run(L, offs, w, x, steps, threads=0)
    rm, _ = rel_errors(L.interior(got), L.interior(want))
    assert rm <= 1e-12, rm


@pytest.mark.parametrize("name", ["box2d1r", "star2d3r", "box2d3r"])
def test_noalloc_matches_plan_twice(gpu, name):
    print(f"case noalloc {name}", flush=True)
    # a grid with several fused launches and a remainder, then a smaller one
    _case(name, [700, 913], 13, 11)
    _case(name, [97, 61], 5, 12)


def test_noalloc_single_steps_and_errors(gpu):
    print("case noalloc single steps / errors", flush=True)
    _case("box2d1r", [130, 257], 4, 13, fusion=1)
    lib = _lib.load()
    L = Layout.with_halo([40, 40], 1)
    Lc = L.to_c()
    w = np.full(9, 1.0 / 9)
    st = _lib.cs_stencil(2, 1, w.ctypes.data_as(_lib.c_double_p))
    fin = ctypes.c_int32()
    opt = _lib.cs_options()
    # null buffers, the same buffer twice, a split precision: refused before any launch
    assert lib.cs_run_2d_noalloc(ctypes.byref(Lc), ctypes.byref(st), None, None, 1, None, None,
                                 ctypes.byref(fin)) == _lib.CS_ERR_INVALID
    dummy = ctypes.c_void_p(4096)
    assert lib.cs_run_2d_noalloc(ctypes.byref(Lc), ctypes.byref(st), dummy, dummy, 1, None, None,
                                 ctypes.byref(fin)) == _lib.CS_ERR_INVALID
    opt.precision = _lib.CS_PREC_BF16X3
    assert lib.cs_run_2d_noalloc(ctypes.byref(Lc), ctypes.byref(st), dummy, ctypes.c_void_p(8192), 1,
                                 ctypes.byref(opt), None, ctypes.byref(fin)) == _lib.CS_ERR_UNSUPPORTED
