"""Host-side guards of the runtime, checked on the CPU through the debug entry points
(include/convstencil_debug.h), plus their GPU counterparts.

* FP16 range guard: the FP16 split forms (CS_PREC_FP16X2 / CS_PREC_FP16W3) hold every
  value as FP16 slices, so a run whose state could reach 2^15 in magnitude is refused
  before anything is launched; one step grows the bound by at most sum |w|.
* Batch error latch: a grid-wide barrier that timed out leaves the batch's grids
  undefined; from the first read of the sticky device word on, every call fails.
"""
import ctypes
import math

import numpy as np
import pytest

import convstencil_amd as cs
from convstencil_amd import _lib, Layout, Plan, Batch, shape_offsets, dense_weights
from convstencil_amd._lib import ConvStencilError


def _dbg():
    lib = _lib.load()
    lib.cs_debug_fp16_range.restype = ctypes.c_int32
    lib.cs_debug_fp16_range.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_double)]
    lib.cs_debug_abs_max.restype = ctypes.c_double
    lib.cs_debug_abs_max.argtypes = [_lib.c_double_p, ctypes.c_int64]
    lib.cs_debug_batch_latch.restype = ctypes.c_int32
    lib.cs_debug_batch_latch.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int32)]
    return lib


def _range(prec, amax, wabs, steps):
    b = ctypes.c_double(-1)
    ok = _dbg().cs_debug_fp16_range(prec, amax, wabs, steps, ctypes.byref(b))
    return ok, b.value


@pytest.mark.parametrize("prec", [_lib.CS_PREC_FP16X2, _lib.CS_PREC_FP16W3])
def test_fp16_range_guard_bounds(prec):
    # normalised positive weights never grow the state: any number of steps is allowed
    assert _range(prec, 1.0, 1.0, 10 ** 9) == (1, 1.0)
    assert _range(prec, 32767.0, 1.0, 100) == (1, 32767.0)
    # at or above 2^15 from the start: refused, whatever the step count
    assert _range(prec, 32768.0, 1.0, 1)[0] == 0
    assert _range(prec, 65536.0, 0.5, 1)[0] == 0
    # growth: sum |w| = 28 (the reference's 1-D default weights 1..7), from |u| <= 10
    ok, b = _range(prec, 10.0, 28.0, 2)
    assert ok == 1 and b == pytest.approx(10.0 * 28.0 ** 2)
    ok, b = _range(prec, 10.0, 28.0, 3)
    assert ok == 0 and b == pytest.approx(10.0 * 28.0 ** 3)
    # the bound overflows to inf for long runs and is refused
    ok, b = _range(prec, 1.0, 2.0, 5000)
    assert ok == 0 and math.isinf(b)
    # a non-finite state is refused
    assert _range(prec, math.inf, 1.0, 1)[0] == 0
    # a zero state (a new plan) may run
    assert _range(prec, 0.0, 28.0, 10 ** 6) == (1, 0.0)


@pytest.mark.parametrize("prec", [_lib.CS_PREC_FP64, _lib.CS_PREC_BF16X3, _lib.CS_PREC_FP32X2])
def test_fp16_range_guard_other_precisions_pass(prec):
    assert _range(prec, 1e300, 28.0, 100)[0] == 1


def test_abs_max_of_upload():
    lib = _dbg()
    a = np.array([0.5, -70000.0, 3.0])
    assert lib.cs_debug_abs_max(a.ctypes.data_as(_lib.c_double_p), a.size) == 70000.0
    a = np.array([0.5, np.nan, 3.0])
    assert math.isinf(lib.cs_debug_abs_max(a.ctypes.data_as(_lib.c_double_p), a.size))
    a = np.array([np.inf])
    assert math.isinf(lib.cs_debug_abs_max(a.ctypes.data_as(_lib.c_double_p), a.size))
    a = np.zeros(0)
    assert lib.cs_debug_abs_max(a.ctypes.data_as(_lib.c_double_p), 0) == 0.0


def test_batch_error_latch_is_sticky():
    lib = _dbg()
    flags = (ctypes.c_uint32 * 6)(0, 0, 1, 0, 0, 0)
    st = (ctypes.c_int32 * 6)()
    assert lib.cs_debug_batch_latch(flags, 6, st) == 0
    assert list(st) == [_lib.CS_OK, _lib.CS_OK] + [_lib.CS_ERR_DEVICE] * 4
    flags = (ctypes.c_uint32 * 3)(0, 0, 0)
    st = (ctypes.c_int32 * 3)()
    lib.cs_debug_batch_latch(flags, 3, st)
    assert list(st) == [_lib.CS_OK] * 3


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp16x2", "fp16w3"])
def test_gpu_plan_refuses_fp16_overflow(gpu, prec):
    L = Layout.with_halo([64, 80], 3)
    offs = shape_offsets("box2d3r")
    w = cs.fill_weights_host(len(offs), 5)
    dense = dense_weights(2, 3, offs, w)
    x = np.zeros(L.cells)
    L.interior(x)[...] = 1.0
    L.interior(x)[0, 10, 10] = 65536.0  # 2^16: overflows an FP16 slice
    with Plan(L, 3, dense, precision=prec) as p:
        p.upload(x)
        with pytest.raises(ConvStencilError) as e:
            p.run(1)
        assert e.value.status == _lib.CS_ERR_INVALID
    # the same grid runs in BF16x3 (no range limit) and after scaling into range in FP16
    with Plan(L, 3, dense, precision="bf16x3") as p:
        p.upload(x)
        p.run(1)
    with Plan(L, 3, dense, precision=prec) as p:
        p.upload(x / 4)
        p.run(4)
    # growth: unnormalised weights (sum 49) from |u| <= 1 reach 2^15 within 3 steps
    with Plan(L, 3, dense * 49 / np.sum(dense), precision=prec) as p:
        p.fill_random(3)
        p.run(2)  # bound 49^2 = 2401
        with pytest.raises(ConvStencilError):
            p.run(1)  # bound 49^3 > 2^15


@pytest.mark.gpu
def test_gpu_batch_refuses_fp16_overflow(gpu):
    Ls = [Layout.with_halo([70, 90], 1), Layout.with_halo([130, 64], 3)]
    shapes = ["box2d1r", "star2d3r"]
    dens = []
    for i, (L, sh) in enumerate(zip(Ls, shapes)):
        offs = shape_offsets(sh)
        dens.append(dense_weights(2, 1 if sh == "box2d1r" else 3, offs, cs.fill_weights_host(len(offs), 20 + i)))
    with Batch(Ls, [1, 3], dens, precision="fp16x2") as b:
        b.fill_random([1, 2])
        b.run(3)
        x = np.zeros(Ls[1].cells)
        Ls[1].interior(x)[0, 5, 5] = 40000.0
        b.upload(1, x)
        with pytest.raises(ConvStencilError) as e:
            b.run(1)
        assert e.value.status == _lib.CS_ERR_INVALID
