"""The --breakdown ladder (reference src/2d/gpu.cu:434-801, src/2d/main.cu:349-357) on the GPU.

Every rung computes the same result as the main launcher: the reference-named rung
launchers against gpu_box_2d1r on the reference layout with a random halo, the rung
kernels forced into plans against the CPU oracle, and the command line's report.
"""
import os
import subprocess

import numpy as np
import pytest

from convstencil_amd import Plan, Layout, shape_offsets, dense_weights, SHAPES
from convstencil_amd import _lib
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _launch(name, x, w, times, m, n):
    lib = _lib.load()
    out = np.zeros_like(x)
    rc = getattr(lib, name)(x.ctypes.data_as(_lib.c_double_p), out.ctypes.data_as(_lib.c_double_p),
                            w.ctypes.data_as(_lib.c_double_p), times, m, n)
    assert rc == 0
    return out


@pytest.mark.parametrize("m,n,times", [(77, 131, 3), (64, 64, 2), (130, 47, 5)])
def test_rungs_match_main_launcher(gpu, m, n, times):
    print(f"case {m}x{n} times={times}", flush=True)
    rng = np.random.default_rng(m * n + times)
    x = rng.integers(0, 100, (m + 6) * (n + 8)).astype(np.float64)  # the reference's fill, halo included
    w = rng.uniform(0, 1, 49) / 49
    main = _launch("gpu_box_2d1r", x, w, times, m, n)
    L = Layout(2, [m + 6, n + 8], [3, 4], [m, n])
    for k in range(1, 5):
        got = _launch(f"gpu_box_2d1r_breakdown{k}", x, w, times, m, n)
        rm, _ = rel_errors(L.interior(got), L.interior(main))
        assert rm <= 1e-12, (k, rm)
        # halo cells come back as the reference's ping-pong leaves them
        assert np.array_equal(got.reshape(L.ext)[~L.interior_mask().reshape(L.ext)],
                              main.reshape(L.ext)[~L.interior_mask().reshape(L.ext)]), k


@pytest.mark.parametrize("vid,name", [(911000, "box2d1r"), (913000, "box2d3r"), (911010, "star2d1r"),
                                      (913010, "star2d3r"), (921000, "box2d1r"), (923000, "box2d3r"),
                                      (921010, "star2d1r"), (923010, "star2d3r"), (933101, "box2d3r"),
                                      (931101, "box2d1r"), (933111, "star2d3r"), (931111, "star2d1r")])
def test_rung_kernels_against_oracle(gpu, vid, name):
    print(f"case {vid} {name}", flush=True)
    ndim, r, kind = SHAPES[name]
    for sizes in ([70, 150], [200, 131]):
        L = Layout.with_halo(sizes, r)
        offs = shape_offsets(name)
        w = oracle.fill_weights(len(offs), vid % 97)
        x = oracle.fill_data(L, vid % 89)
        with Plan(L, r, dense_weights(2, r, offs, w), variant=vid) as p:
            assert f"variant={vid}" in p.describe()
            p.upload(x)
            p.run(4)
            got = p.download()
        want = oracle.run(L, offs, w, x, 4)
        rm, _ = rel_errors(L.interior(got), L.interior(want))
        assert rm <= 1e-12, (sizes, rm)
        assert np.array_equal(got[~L.interior_mask()], want[~L.interior_mask()])


def test_cli_breakdown_report(gpu):
    exe = os.path.join(ROOT, "convstencil_amd", "bin", "convstencil_2d")
    p = subprocess.run([exe, "box2d1r", "256", "256", "3", "--breakdown"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    pos = [p.stdout.find(f"Experiment - Breakdown(2D) {k}: \nTime = ") for k in range(1, 5)]
    assert all(q >= 0 for q in pos) and pos == sorted(pos), p.stdout
    assert p.stdout.find("ConvStencil(2D): \nTime = ") > pos[-1]
    assert p.stdout.count("GStencil/s = ") == 5
