"""Python host API over libconvstencil (ctypes).  No torch on this path.

`Plan` wraps cs_plan_* (device-resident ping-pong grids); `run` is the one-shot
host-to-host call; the reference-named functions at the bottom mirror the
reference's host launchers (src/2d/2d_utils.h:77-79, src/1d/1d_utils.h:75,85,
src/3d/3d_utils.h:75-77) with numpy arrays for the host pointers.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check
from .shapes import Layout

PRECISIONS = {"fp64": _lib.CS_PREC_FP64, "bf16x3": _lib.CS_PREC_BF16X3, "fp32x2": _lib.CS_PREC_FP32X2,
              "fp16x2": _lib.CS_PREC_FP16X2, "fp16w3": _lib.CS_PREC_FP16W3}


def _dptr(a):
    return a.ctypes.data_as(_lib.c_double_p)


def _contig(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if a.size < n:
        raise ValueError(f"array has {a.size} elements, need {n}")
    return a


class Report:
    def __init__(self, r):
        self.elapsed_ms = r.elapsed_ms
        self.gstencil_per_s = r.gstencil_per_s
        self.launches = r.launches
        self.fusion_depth = r.fusion_depth
        self.variant = r.variant

    def __repr__(self):
        return (f"Report(elapsed_ms={self.elapsed_ms:.4f}, gstencil_per_s={self.gstencil_per_s:.2f}, "
                f"launches={self.launches}, fusion_depth={self.fusion_depth}, variant={self.variant})")


def _options(precision="fp64", fusion_depth=0, use_graph=False, variant=0, autotune=0, verbose=0):
    o = _lib.cs_options()
    o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.fusion_depth = int(fusion_depth)
    o.use_graph = int(bool(use_graph))
    o.tile_variant = int(variant)
    o.verbose = int(verbose)
    o.autotune = int(autotune)
    return o


class Plan:
    """A stencil on a padded grid held in two device buffers."""

    def __init__(self, layout: Layout, radius: int, weights_dense, **opts):
        lib = _lib.load()
        self.layout = layout
        self.radius = int(radius)
        self._w = np.ascontiguousarray(weights_dense, dtype=np.float64).reshape(-1)
        need = (2 * self.radius + 1) ** layout.ndim
        if self._w.size != need:
            raise ValueError(f"need {need} dense weights, got {self._w.size}")
        st = _lib.cs_stencil(layout.ndim, self.radius, _dptr(self._w))
        self._L = layout.to_c()
        self._opt = _options(**opts)
        h = ctypes.c_void_p()
        check(lib.cs_plan_create(ctypes.byref(h), ctypes.byref(self._L), ctypes.byref(st), ctypes.byref(self._opt)),
              "cs_plan_create")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cs_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, host):
        a = _contig(host, self.layout.cells)
        check(self._lib.cs_plan_upload(self._h, _dptr(a)), "cs_plan_upload")

    def fill_random(self, seed):
        check(self._lib.cs_plan_fill_random(self._h, ctypes.c_uint64(seed)), "cs_plan_fill_random")

    def run(self, steps):
        r = _lib.cs_report()
        check(self._lib.cs_plan_run(self._h, int(steps), ctypes.byref(r)), "cs_plan_run")
        return Report(r)

    def time(self, steps, warmup=1, reps=3):
        r = _lib.cs_report()
        check(self._lib.cs_plan_time(self._h, int(steps), int(warmup), int(reps), ctypes.byref(r)), "cs_plan_time")
        return Report(r)

    def sync(self):
        check(self._lib.cs_plan_sync(self._h), "cs_plan_sync")

    def download(self):
        out = np.empty(self.layout.cells, dtype=np.float64)
        check(self._lib.cs_plan_download(self._h, _dptr(out)), "cs_plan_download")
        return out

    def device_buffers(self):
        cur, oth = _lib.c_double_p(), _lib.c_double_p()
        check(self._lib.cs_plan_device_buffers(self._h, ctypes.byref(cur), ctypes.byref(oth)), "device_buffers")
        return ctypes.cast(cur, ctypes.c_void_p).value, ctypes.cast(oth, ctypes.c_void_p).value

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        check(self._lib.cs_plan_describe(self._h, buf, 512), "cs_plan_describe")
        return buf.value.decode()


class Batch:
    """Many independent 2-D grids advanced together by one persistent launch per run
    (cs_batch_*): each grid has its own layout, radius and weights."""

    def __init__(self, layouts, radii, weights_dense, precision="fp64", use_graph=False):
        lib = _lib.load()
        n = len(layouts)
        if n == 0 or len(radii) != n or len(weights_dense) != n:
            raise ValueError("need one radius and one weight array per layout")
        self.layouts = list(layouts)
        self._w = [np.ascontiguousarray(w, dtype=np.float64).reshape(-1) for w in weights_dense]
        Ls = (_lib.cs_layout * n)(*[L.to_c() for L in self.layouts])
        Ss = (_lib.cs_stencil * n)(*[_lib.cs_stencil(2, int(r), _dptr(w)) for r, w in zip(radii, self._w)])
        self._opt = _options(precision=precision, use_graph=use_graph)
        h = ctypes.c_void_p()
        check(lib.cs_batch_create(ctypes.byref(h), n, Ls, Ss, ctypes.byref(self._opt)), "cs_batch_create")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, i, host):
        a = _contig(host, self.layouts[i].cells)
        check(self._lib.cs_batch_upload(self._h, int(i), _dptr(a)), "cs_batch_upload")

    def fill_random(self, seeds):
        s = (ctypes.c_uint64 * len(self.layouts))(*[int(v) for v in seeds])
        check(self._lib.cs_batch_fill_random(self._h, s), "cs_batch_fill_random")

    def run(self, steps):
        r = _lib.cs_report()
        check(self._lib.cs_batch_run(self._h, int(steps), ctypes.byref(r)), "cs_batch_run")
        return Report(r)

    def time(self, steps, warmup=1, reps=3):
        r = _lib.cs_report()
        check(self._lib.cs_batch_time(self._h, int(steps), int(warmup), int(reps), ctypes.byref(r)), "cs_batch_time")
        return Report(r)

    def download(self, i):
        out = np.empty(self.layouts[i].cells, dtype=np.float64)
        check(self._lib.cs_batch_download(self._h, int(i), _dptr(out)), "cs_batch_download")
        return out

    def describe(self):
        buf = ctypes.create_string_buffer(512)
        check(self._lib.cs_batch_describe(self._h, buf, 512), "cs_batch_describe")
        return buf.value.decode()


def run(layout: Layout, radius, weights_dense, host_in, steps, **opts):
    """One-shot host-to-host run; returns (padded output, Report)."""
    lib = _lib.load()
    w = np.ascontiguousarray(weights_dense, dtype=np.float64).reshape(-1)
    st = _lib.cs_stencil(layout.ndim, int(radius), _dptr(w))
    a = _contig(host_in, layout.cells)
    out = np.empty(layout.cells, dtype=np.float64)
    L = layout.to_c()
    o = _options(**opts)
    r = _lib.cs_report()
    check(lib.cs_run(ctypes.byref(L), ctypes.byref(st), _dptr(a), _dptr(out), int(steps), ctypes.byref(o),
                     ctypes.byref(r)), "cs_run")
    return out, Report(r)


def fill_random_host(layout: Layout, seed):
    a = np.empty(layout.cells, dtype=np.float64)
    L = layout.to_c()
    _lib.load().cs_fill_random_host(_dptr(a), ctypes.byref(L), ctypes.c_uint64(seed))
    return a


def fill_weights_host(count, seed):
    w = np.empty(int(count), dtype=np.float64)
    _lib.load().cs_fill_weights_host(_dptr(w), int(count), ctypes.c_uint64(seed))
    return w


def save_txt(path, layout: Layout, arr):
    """The reference's text dump of a padded array (save_to_txt, src/2d/main.cu:36-55 and
    the 1-D / 3-D variants): "%.0f" values, as its drivers write output_gpu.txt."""
    a = _contig(arr, layout.cells)
    L = layout.to_c()
    check(_lib.load().cs_grid_save_txt(os.fsencode(path), ctypes.byref(L), _dptr(a)), "cs_grid_save_txt")


def save_f64(path, layout: Layout, arr):
    """Exact binary grid file: "CSGRID64" header with the layout, then the padded array."""
    a = _contig(arr, layout.cells)
    L = layout.to_c()
    check(_lib.load().cs_grid_save_f64(os.fsencode(path), ctypes.byref(L), _dptr(a)), "cs_grid_save_f64")


def load_f64(path):
    """Reads a file written by save_f64 (or cs_grid_save_f64); returns (Layout, flat array)."""
    lib = _lib.load()
    L = _lib.cs_layout()
    check(lib.cs_grid_load_f64(os.fsencode(path), ctypes.byref(L), None, 0), "cs_grid_load_f64")
    layout = Layout(3, list(L.ext), list(L.org), list(L.n))
    layout.ndim = int(L.ndim)
    out = np.empty(layout.cells, dtype=np.float64)
    check(lib.cs_grid_load_f64(os.fsencode(path), ctypes.byref(L), _dptr(out), out.size), "cs_grid_load_f64")
    return layout, out


def device_count():
    c = ctypes.c_int32(0)
    check(_lib.load().cs_device_count(ctypes.byref(c)), "cs_device_count")
    return c.value


# ---- mirrors of the reference's host launchers (same argument meaning) ----

def _launcher(name, in_arr, out_arr, params, nparams, times, sizes, cells):
    lib = _lib.load()
    a = _contig(in_arr, cells)
    if out_arr.dtype != np.float64 or not out_arr.flags.c_contiguous or out_arr.size < cells:
        raise ValueError("out must be a contiguous float64 array of the padded size")
    p = _contig(params, nparams)
    rc = getattr(lib, name)(_dptr(a), _dptr(out_arr), _dptr(p), int(times), *[int(s) for s in sizes])
    check(rc, name)


def gpu_1d1r(in_arr, out_arr, params, times, input_n):
    _launcher("gpu_1d1r", in_arr, out_arr, params, 7, times, [input_n], input_n + 7)


def gpu_star_1d2r_step2(in_arr, out_arr, params, times, input_n):
    _launcher("gpu_star_1d2r_step2", in_arr, out_arr, params, 9, times, [input_n], input_n + 8)


def gpu_box_2d1r(in_arr, out_arr, params, times, input_m, input_n):
    _launcher("gpu_box_2d1r", in_arr, out_arr, params, 49, times, [input_m, input_n], (input_m + 6) * (input_n + 8))


def gpu_box_2d3r(in_arr, out_arr, params, times, input_m, input_n):
    _launcher("gpu_box_2d3r", in_arr, out_arr, params, 49, times, [input_m, input_n], (input_m + 6) * (input_n + 8))


def gpu_box_3d1r(in_arr, out_arr, params, times, input_h, input_m, input_n):
    _launcher("gpu_box_3d1r", in_arr, out_arr, params, 196, times, [input_h, input_m, input_n],
              (input_h + 6) * (input_m + 6) * (input_n + 6))


def gpu_star_3d1r(in_arr, out_arr, params, times, input_h, input_m, input_n):
    _launcher("gpu_star_3d1r", in_arr, out_arr, params, 44, times, [input_h, input_m, input_n],
              (input_h + 6) * (input_m + 6) * (input_n + 6))
