"""ctypes binding of libconvstencil.so (the C ABI declared in include/convstencil.h).

The library is built in-tree (``__graft_entry__.build()`` or ``make -C
convstencil_amd/csrc``) into ``convstencil_amd/lib/libconvstencil.so``.  There is
no fallback: if the library is missing, loading raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libconvstencil.so")

CS_OK, CS_ERR_INVALID, CS_ERR_UNSUPPORTED, CS_ERR_ALLOC, CS_ERR_DEVICE, CS_ERR_NO_DEVICE = range(6)
CS_PREC_FP64, CS_PREC_BF16X3, CS_PREC_FP32X2, CS_PREC_FP16X2, CS_PREC_FP16W3 = range(5)

c_double_p = ctypes.POINTER(ctypes.c_double)


class cs_layout(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("reserved", ctypes.c_int32), ("ext", ctypes.c_int64 * 3),
                ("org", ctypes.c_int64 * 3), ("n", ctypes.c_int64 * 3)]


class cs_stencil(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("radius", ctypes.c_int32), ("weights", c_double_p)]


class cs_options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("fusion_depth", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("tile_variant", ctypes.c_int32), ("verbose", ctypes.c_int32), ("autotune", ctypes.c_int32)]


class cs_report(ctypes.Structure):
    _fields_ = [("elapsed_ms", ctypes.c_double), ("gstencil_per_s", ctypes.c_double), ("launches", ctypes.c_int64),
                ("fusion_depth", ctypes.c_int32), ("variant", ctypes.c_int32)]


# name -> (restype, argtypes); every function declared in include/convstencil.h
_P = ctypes.c_void_p
SIGNATURES = {
    "cs_version": (ctypes.c_char_p, []),
    "cs_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "cs_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    "cs_plan_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(cs_layout), ctypes.POINTER(cs_stencil),
                                      ctypes.POINTER(cs_options)]),
    "cs_plan_destroy": (ctypes.c_int, [_P]),
    "cs_plan_upload": (ctypes.c_int, [_P, c_double_p]),
    "cs_plan_fill_random": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "cs_plan_run": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(cs_report)]),
    "cs_plan_time": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(cs_report)]),
    "cs_plan_sync": (ctypes.c_int, [_P]),
    "cs_plan_download": (ctypes.c_int, [_P, c_double_p]),
    "cs_plan_device_buffers": (ctypes.c_int, [_P, ctypes.POINTER(c_double_p), ctypes.POINTER(c_double_p)]),
    "cs_plan_describe": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int32]),
    "cs_run": (ctypes.c_int, [ctypes.POINTER(cs_layout), ctypes.POINTER(cs_stencil), c_double_p, c_double_p,
                              ctypes.c_int64, ctypes.POINTER(cs_options), ctypes.POINTER(cs_report)]),
    "cs_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.POINTER(cs_layout),
                                       ctypes.POINTER(cs_stencil), ctypes.POINTER(cs_options)]),
    "cs_batch_destroy": (ctypes.c_int, [_P]),
    "cs_batch_upload": (ctypes.c_int, [_P, ctypes.c_int32, c_double_p]),
    "cs_batch_fill_random": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "cs_batch_run": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(cs_report)]),
    "cs_batch_time": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(cs_report)]),
    "cs_batch_download": (ctypes.c_int, [_P, ctypes.c_int32, c_double_p]),
    "cs_batch_describe": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int32]),
    "cs_tune_cache_path": (ctypes.c_char_p, []),
    "cs_tune_cache_reset": (ctypes.c_int, []),
    "cs_tune_lookup": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
    "cs_tune_store": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_double]),
    "cs_tune_key": (ctypes.c_int, [ctypes.POINTER(cs_layout), ctypes.POINTER(cs_stencil), ctypes.POINTER(cs_options),
                                   ctypes.c_char_p, ctypes.c_int32]),
    "cs_run_2d_noalloc": (ctypes.c_int, [ctypes.POINTER(cs_layout), ctypes.POINTER(cs_stencil), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(cs_options), ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int32)]),
    "cs_fill_random_host": (None, [c_double_p, ctypes.POINTER(cs_layout), ctypes.c_uint64]),
    "cs_fill_weights_host": (None, [c_double_p, ctypes.c_int32, ctypes.c_uint64]),
    "cs_grid_save_txt": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(cs_layout), c_double_p]),
    "cs_grid_save_f64": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(cs_layout), c_double_p]),
    "cs_grid_load_f64": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(cs_layout), c_double_p, ctypes.c_int64]),
    "gpu_1d1r": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int]),
    "gpu_star_1d2r_step2": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int]),
    "gpu_box_2d1r": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gpu_box_2d3r": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gpu_box_2d1r_breakdown1": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int]),
    "gpu_box_2d1r_breakdown2": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int]),
    "gpu_box_2d1r_breakdown3": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int]),
    "gpu_box_2d1r_breakdown4": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int]),
    "gpu_box_3d1r": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int]),
    **{f"gpu_1d1r_breakdown{k}": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int])
       for k in range(1, 5)},
    **{f"gpu_box_3d1r_breakdown{k}": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int]) for k in range(1, 5)},
    "gpu_star_3d1r": (ctypes.c_int, [c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]),
}

_lib = None


class ConvStencilError(RuntimeError):
    def __init__(self, what, status):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


def load():
    """Load libconvstencil.so and set its prototypes.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libconvstencil.so not built at {LIB_PATH}; run __graft_entry__.build() "
                          f"or make -C convstencil_amd/csrc")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def status_string(status):
    return load().cs_status_string(int(status)).decode()


def check(status, what):
    if status != CS_OK:
        raise ConvStencilError(what, status)
