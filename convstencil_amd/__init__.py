"""ConvStencil for AMD Instinct MI355X (gfx950).

Host side of a C++/HIP library (libconvstencil.so, C ABI in include/convstencil.h)
that runs the ConvStencil time loop -- stencil2row on FP64 MFMA, temporal fusion
in LDS, hipGraph-captured time loops -- on one GPU.
"""
from .shapes import SHAPES, Layout, dense_weights, offsets, shape_offsets  # noqa: F401
from .api import (Batch, Plan, Report, run, fill_random_host, fill_weights_host, device_count,  # noqa: F401
                  save_txt, save_f64, load_f64,
                  gpu_1d1r, gpu_star_1d2r_step2, gpu_box_2d1r, gpu_box_2d3r, gpu_box_3d1r, gpu_star_3d1r)

__version__ = "0.1.0"
