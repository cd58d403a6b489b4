"""convstencil_{1d,2d,3d} command lines -- CPU only (no GPU needed).

The grammar, help text, prompts, messages and exit codes are the reference's
(src/1d/main.cu:49-100, src/2d/main.cu:117-219, src/3d/main.cu:469-563).  On a
host without a GPU a well-formed run fails at plan creation with exit code 1.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from convstencil_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "convstencil_amd", "bin")


def run(prog, *args, stdin=""):
    p = subprocess.run([os.path.join(BIN, prog), *args], input=stdin, capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("prog,args", [("convstencil_1d", ["1d1r", "10"]), ("convstencil_2d", ["box2d1r", "1", "2"]),
                                       ("convstencil_3d", ["box3d1r", "1", "2", "3"])])
def test_too_few_args_prints_help(prog, args):
    rc, out, err = run(prog, *args)
    assert rc == 1
    assert out.startswith("Program name: convstencil_")
    assert "--custom" in out


@pytest.mark.parametrize("prog,args", [("convstencil_1d", ["1d3r", "10", "2"]),
                                       ("convstencil_2d", ["box2d2r", "8", "8", "2"]),
                                       ("convstencil_3d", ["box3d3r", "8", "8", "8", "2"])])
def test_bad_shape_prints_help(prog, args):
    rc, out, err = run(prog, *args)
    assert rc == 1 and "Usage:" in out


def test_help_text_matches_reference_quirks():
    assert "Usage: convstencil_2 shape input_size time_iteration_size" in run("convstencil_1d")[1]
    assert run("convstencil_3d")[1].startswith("Program name: convstencil_2d\n")


@pytest.mark.parametrize("prog,args", [("convstencil_2d", ["box2d1r", "abc", "8", "2"]),
                                       ("convstencil_1d", ["1d1r", "x", "2"]),
                                       ("convstencil_3d", ["star3d1r", "8", "8", "q", "2"])])
def test_non_integer(prog, args):
    rc, out, err = run(prog, *args)
    assert rc == 1
    assert err == "Invalid argument: cannot convert the parameter(s) to integer.\n"


def test_out_of_range():
    rc, out, err = run("convstencil_2d", "box2d1r", "99999999999", "8", "2")
    assert rc == 1
    assert err == "Argument out of range: the parameter(s) is(are) too large.\n"


@pytest.mark.parametrize("prog,args,count", [("convstencil_2d", ["box2d1r", "8", "8", "2"], 9),
                                             ("convstencil_2d", ["star2d1r", "8", "8", "2"], 5),
                                             ("convstencil_1d", ["1d1r", "64", "2"], 3),
                                             ("convstencil_1d", ["1d2r", "64", "2"], 5),
                                             ("convstencil_3d", ["box3d1r", "4", "4", "4", "2"], 27),
                                             ("convstencil_3d", ["star3d1r", "4", "4", "4", "2"], 7)])
def test_custom_short_stdin_exits_1(prog, args, count):
    rc, out, err = run(prog, *args, "--custom", stdin="0.5 " * (count - 1))
    assert rc == 1
    assert out == f"Please enter {count} parameters:\n"


def test_no_gpu_exit_code_and_info_line():
    rc, out, err = run("convstencil_2d", "box2d1r", "64", "64", "3")
    if "no gfx950 device" not in err:
        pytest.skip("a GPU is visible; covered by the GPU tests")
    assert rc == 1
    assert "INFO: shape = box_2d1r, m = 64, n = 64, times = 3" in out


def _fused(exact, p):
    lib = _lib.load()
    lib.cs_debug_fused_2d.restype = None
    lib.cs_debug_fused_2d.argtypes = [ctypes.c_int32, _lib.c_double_p, _lib.c_double_p]
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(49)
    lib.cs_debug_fused_2d(int(exact), p.ctypes.data_as(_lib.c_double_p), out.ctypes.data_as(_lib.c_double_p))
    return out


def _cube(p):
    k = np.asarray(p, dtype=np.float64).reshape(3, 3)
    two = np.zeros((5, 5))
    for a in range(3):
        for b in range(3):
            two[a:a + 3, b:b + 3] += k[a, b] * k
    three = np.zeros((7, 7))
    for a in range(5):
        for b in range(5):
            three[a:a + 3, b:b + 3] += two[a, b] * k
    return three.reshape(-1)


def test_exact_fusion_is_the_cube():
    rng = np.random.default_rng(3)
    for _ in range(10):
        p = rng.uniform(-1, 1, 9)
        np.testing.assert_allclose(_fused(1, p), _cube(p), rtol=0, atol=1e-14)


def test_reference_table_and_its_nine_mistyped_entries():
    # equal weights: the reference's table equals the exact cube everywhere
    p = np.full(9, 0.3)
    np.testing.assert_allclose(_fused(0, p), _cube(p), rtol=1e-15, atol=0)
    # integer weights (exact arithmetic): 40 entries agree, exactly these 9 differ
    p = np.arange(1, 10, dtype=np.float64)
    ref, exact = _fused(0, p), _cube(p)
    differ = sorted(int(i) for i in np.nonzero(ref != exact)[0])
    assert differ == [18, 19, 30, 31, 32, 33, 34, 35, 41]
    # spot values of the reference's formulas (src/2d/main.cu:249,257): entry 0 = p0^3,
    # entry 24 (centre) includes pow(p4, 3)
    assert ref[0] == 1.0 and ref[48] == 729.0
    assert ref[24] == exact[24]
