"""The --breakdown ladders on the GPU: 2-D (reference src/2d/gpu.cu:434-801,
src/2d/main.cu:349-357), 1-D (src/1d/gpu_1r.cu:305-651, src/1d/main.cu:224-228) and 3-D box
(src/3d/gpu_box.cu:538-904, src/3d/main.cu:623-628).

Every rung computes the same result as the main launcher: the reference-named rung
launchers against the main launcher on the reference layout with a random halo, the rung
kernels forced into plans against the CPU oracle, and the command lines' reports.
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


# ---------------------------------------------------------------- 1-D and 3-D ladders

def _launch_nd(name, x, w, times, *sizes):
    lib = _lib.load()
    out = np.zeros_like(x)
    rc = getattr(lib, name)(x.ctypes.data_as(_lib.c_double_p), out.ctypes.data_as(_lib.c_double_p),
                            w.ctypes.data_as(_lib.c_double_p), times, *sizes)
    assert rc == 0
    return out


@pytest.mark.parametrize("n,times", [(1000, 3), (4097, 2), (64, 5)])
def test_rungs_match_main_launcher_1d(gpu, n, times):
    print(f"case 1d n={n} times={times}", flush=True)
    rng = np.random.default_rng(n + times)
    x = rng.integers(0, 10000, n + 7).astype(np.float64)  # the reference's fill, halo included
    w = rng.uniform(0, 1, 7) / 7
    main = _launch_nd("gpu_1d1r", x, w, times, n)
    for k in range(1, 5):
        got = _launch_nd(f"gpu_1d1r_breakdown{k}", x, w, times, n)
        rm, _ = rel_errors(got, main)
        assert rm <= 1e-12, (k, rm)
        # out[j] = final[j + 1]: the halo cells (out[0:3], out[n + 3:n + 6]) come back unchanged
        assert np.array_equal(got[:3], main[:3]) and np.array_equal(got[n + 3:n + 6], main[n + 3:n + 6]), k


@pytest.mark.parametrize("h,m,n,times", [(9, 12, 70, 2), (16, 16, 64, 3)])
def test_rungs_match_main_launcher_3d(gpu, h, m, n, times):
    print(f"case 3d {h}x{m}x{n} times={times}", flush=True)
    rng = np.random.default_rng(h * m * n + times)
    x = rng.integers(0, 100, (h + 6) * (m + 6) * (n + 6)).astype(np.float64)
    w = rng.uniform(0, 1, 196) / 343
    main = _launch_nd("gpu_box_3d1r", x, w, times, h, m, n)
    L = Layout(3, [h + 6, m + 6, n + 6], [3, 3, 3], [h, m, n])
    for k in range(1, 5):
        got = _launch_nd(f"gpu_box_3d1r_breakdown{k}", x, w, times, h, m, n)
        rm, _ = rel_errors(L.interior(got), L.interior(main))
        assert rm <= 1e-12, (k, rm)
        assert np.array_equal(got[~L.interior_mask()], main[~L.interior_mask()]), k


def _offsets(ndim, r, kind):
    rng3 = range(-r, r + 1)
    if ndim == 1:
        return [(0, 0, d) for d in rng3]
    # the star class: taps on the axes only
    return [(a, b, c) for a in rng3 for b in rng3 for c in rng3 if kind == "box" or (a != 0) + (b != 0) + (c != 0) <= 1]


@pytest.mark.parametrize("ndim,vid,r,kind", [
    (1, 911000, 1, "box"), (1, 913000, 3, "box"), (1, 921000, 1, "box"), (1, 923000, 3, "box"),
    (1, 933001, 3, "box"),
    (3, 911000, 1, "box"), (3, 911010, 1, "star"), (3, 913000, 3, "box"), (3, 913010, 3, "star"),
    (3, 921000, 1, "box"), (3, 921010, 1, "star"), (3, 923000, 3, "box"), (3, 923010, 3, "star"),
    (3, 933001, 3, "box"), (3, 933011, 3, "star")])
def test_rung_kernels_against_oracle_nd(gpu, ndim, vid, r, kind):
    print(f"case {ndim}d {vid} r{r} {kind}", flush=True)
    offs = _offsets(ndim, r, kind)
    size_sets = ([3000], [129]) if ndim == 1 else ([11, 13, 70], [20, 9, 33])
    for sizes in size_sets:
        L = Layout.with_halo(sizes, r)
        w = oracle.fill_weights(len(offs), vid % 97)
        x = oracle.fill_data(L, vid % 89)
        with Plan(L, r, dense_weights(ndim, r, offs, w), variant=vid) as p:
            assert f"variant={vid}" in p.describe()
            p.upload(x)
            p.run(3)
            got = p.download()
        want = oracle.run(L, offs, w, x, 3)
        rm, _ = rel_errors(L.interior(got), L.interior(want))
        assert rm <= 1e-12, (sizes, rm)
        assert np.array_equal(got[~L.interior_mask()], want[~L.interior_mask()])


@pytest.mark.parametrize("prog,args,dim", [("convstencil_1d", ["1d1r", "4096", "3"], "1D"),
                                           ("convstencil_3d", ["box3d1r", "16", "16", "64", "2"], "3D")])
def test_cli_breakdown_report_nd(gpu, prog, args, dim):
    exe = os.path.join(ROOT, "convstencil_amd", "bin", prog)
    p = subprocess.run([exe, *args, "--breakdown"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    pos = [p.stdout.find(f"Experiment - Breakdown({dim}) {k}: \nTime = ") for k in range(1, 5)]
    assert all(q >= 0 for q in pos) and pos == sorted(pos), p.stdout
    assert p.stdout.find(f"ConvStencil({dim}): \nTime = ") > pos[-1]
    assert p.stdout.count("GStencil/s = ") == 5
