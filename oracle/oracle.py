"""ctypes binding of the CPU oracle (oracle/cs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's CPU-baseline leg.  The product never imports it.

The oracle restates the reference's naive CPU loops (src/1d/main.cu:38-47,
src/2d/main.cu:57-115, src/3d/main.cu:42-467) and its ping-pong driver; see the
header of cs_oracle.c for the exact semantics and tests/test_oracle.py for the
checks that pin it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libcs_oracle.so")
_lib = None
_i64x3 = ctypes.c_int64 * 3
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.cso_splitmix64.restype = ctypes.c_uint64
        lib.cso_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        lib.cso_stream_seed.restype = ctypes.c_uint64
        lib.cso_stream_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        lib.cso_fill_data.restype = None
        lib.cso_fill_data.argtypes = [_dp, _i64x3, _i64x3, _i64x3, ctypes.c_uint64]
        lib.cso_fill_weights.restype = None
        lib.cso_fill_weights.argtypes = [_dp, ctypes.c_int, ctypes.c_uint64]
        lib.cso_step.restype = None
        lib.cso_step.argtypes = [_dp, _dp, _i64x3, _i64x3, _i64x3, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), _dp,
                                 ctypes.c_int]
        lib.cso_run.restype = ctypes.c_int
        lib.cso_run.argtypes = [_dp, _dp, _i64x3, _i64x3, _i64x3, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), _dp,
                                ctypes.c_int64, ctypes.c_int]
        lib.cso_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _geo(layout):
    return _i64x3(*layout.ext), _i64x3(*layout.org), _i64x3(*layout.n)


def fill_data(layout, seed):
    a = np.empty(layout.cells, dtype=np.float64)
    load().cso_fill_data(_p(a), *_geo(layout), ctypes.c_uint64(seed))
    return a


def fill_weights(count, seed):
    w = np.empty(int(count), dtype=np.float64)
    load().cso_fill_weights(_p(w), int(count), ctypes.c_uint64(seed))
    return w


def splitmix_stream(seed, stream, count):
    lib = load()
    st = ctypes.c_uint64(lib.cso_stream_seed(ctypes.c_uint64(seed), ctypes.c_uint64(stream)))
    return [lib.cso_splitmix64(ctypes.byref(st)) for _ in range(count)]


def run(layout, offs, w, host_in, steps, second=None, threads=1):
    """Advance `host_in` (padded, flat) by `steps` steps of the tap list (offs, w).

    `second` is the initial content of the second buffer (the reference zeroes
    it); returns the padded array holding the final state."""
    lib = load()
    a = np.array(host_in, dtype=np.float64, copy=True).reshape(-1)
    b = np.zeros_like(a) if second is None else np.array(second, dtype=np.float64, copy=True).reshape(-1)
    o = np.ascontiguousarray(np.asarray(offs, dtype=np.int32).reshape(-1))
    ww = np.ascontiguousarray(w, dtype=np.float64)
    which = lib.cso_run(_p(a), _p(b), *_geo(layout), len(ww), o.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                        _p(ww), int(steps), int(threads))
    return a if which == 0 else b


def dense_to_taps(ndim, radius, dense):
    """Tap list (row-major offsets, non-zero weights) from a dense weight cube."""
    d = 2 * radius + 1
    cube = np.asarray(dense, dtype=np.float64).reshape((d,) * ndim)
    offs, w = [], []
    for idx in np.ndindex(*cube.shape):
        if cube[idx] == 0.0:
            continue
        o = [0, 0, 0]
        for k, i in enumerate(idx):
            o[3 - ndim + k] = i - radius
        offs.append(tuple(o))
        w.append(cube[idx])
    return offs, np.array(w)


def max_threads():
    return load().cso_max_threads()
