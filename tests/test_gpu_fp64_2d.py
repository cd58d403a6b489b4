"""Plain FP64 2-D path on the GPU against the CPU oracle.

Bar (north_star): relative max error <= 1e-12 against the FP64 oracle, the only
difference being summation order.  Covers every 2-D shape, both MFMA shapes,
every fusion depth and its remainders, ragged and tiny grids, the reference's
padded layout with a non-zero halo (alternating-halo semantics), hipGraph vs
stream (bitwise), determinism, and a full-size 10240^2 x 64-step run checked
on windows whose light cone the oracle recomputes exactly.
"""
import numpy as np
import pytest

from convstencil_amd import Plan, Layout, fill_random_host, shape_offsets, dense_weights, SHAPES
from convstencil_amd import api
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu
TOL = 1e-12


def setup(name, sizes, seed, halo=None):
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo(sizes, r if halo is None else halo)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), seed)
    dense = dense_weights(ndim, r, offs, w)
    x = oracle.fill_data(L, seed)
    return L, r, offs, w, dense, x


def check(L, got, want, tol=TOL):
    rel_max, rel_l2 = rel_errors(L.interior(got), L.interior(want))
    assert rel_max <= tol, (rel_max, rel_l2)
    # halo cells are never written
    m = np.ones(L.ext, dtype=bool)
    o, n = L.org, L.n
    m[o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]] = False
    assert np.array_equal(got.reshape(L.ext)[m], want.reshape(L.ext)[m])
    return rel_max


@pytest.mark.parametrize("name", ["box2d1r", "star2d1r", "star2d3r", "box2d3r"])
@pytest.mark.parametrize("sizes", [[64, 64], [77, 131], [5, 16], [130, 47]])
@pytest.mark.parametrize("steps", [1, 2, 5, 7])
def test_parity_default_plan(gpu, name, sizes, steps):
    print(f"case {name} {sizes} T={steps}", flush=True)
    L, r, offs, w, dense, x = setup(name, sizes, 100 + steps)
    with Plan(L, r, dense) as p:
        p.upload(x)
        p.run(steps)
        got = p.download()
    want = oracle.run(L, offs, w, x, steps)
    check(L, got, want)


def _variant_ids():
    import test_kernel_emulation as em
    return [(v["id"], v["name"], v["R"], v["shape"], v["kf"], v["by"], v["bx"]) for v in em.VARIANTS]


@pytest.mark.parametrize("vid,vname,R,shape,kf,by,bx", _variant_ids())
def test_every_variant(gpu, vid, vname, R, shape, kf, by, bx):
    # 3 x 3 tiles: partial tiles at the far edges and a centre tile inside the interior
    print(f"variant {vname}", flush=True)
    sizes = [3 * by - 3, 3 * bx - 5]
    name = {(1, 0): "box2d1r", (1, 1): "star2d1r", (3, 0): "box2d3r", (3, 1): "star2d3r"}.get((R, shape))
    if name is None:  # radius 2 has no named shape: box or star of radius 2
        offs = [(0, dy, dx) for dy in range(-2, 3) for dx in range(-2, 3) if shape == 0 or dy == 0 or dx == 0]
        L = Layout.with_halo(sizes, 2)
        w = oracle.fill_weights(len(offs), vid)
        dense = dense_weights(2, 2, offs, w)
        x = oracle.fill_data(L, vid)
    else:
        L, R, offs, w, dense, x = setup(name, sizes, vid)
    for steps in (kf, 2 * kf + 1):
        with Plan(L, R, dense, variant=vid) as p:
            p.upload(x)
            p.run(steps)
            got = p.download()
        check(L, got, oracle.run(L, offs, w, x, steps))


def test_reference_layout_random_halo(gpu):
    # the reference's own 2-D layout: rows m+6, cols n+8, interior at (3, 4), 49 weights, random halo
    m, n, times = 45, 70, 5
    L = Layout(2, [m + 6, n + 8], [3, 4], [m, n])
    offs = [(0, dy, dx) for dy in range(-3, 4) for dx in range(-3, 4)]
    w = oracle.fill_weights(49, 77)
    x = oracle.fill_data(L, 77)
    rng = np.random.default_rng(1)
    x3 = x.reshape(L.ext)
    halo = np.ones(L.ext, dtype=bool)
    halo[0, 3:3 + m, 4:4 + n] = False
    x3[halo] = rng.integers(0, 100, int(halo.sum())).astype(np.float64)
    out = np.empty(L.cells)
    api.gpu_box_2d3r(x, out, w, times, m, n)
    want = oracle.run(L, offs, w, x, times)
    check(L, out, want)


def test_golden_fixture_ref_layout(gpu):
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref2d_layout_randhalo.npz"))
    L = Layout(3, z["ext"], z["org"], z["n"])
    L2 = Layout(2, L.ext[1:], L.org[1:], L.n[1:])
    out = np.empty(L2.cells)
    api.gpu_box_2d3r(z["x"], out, z["w"], int(z["steps"]), L2.n[1], L2.n[2])
    check(L2, out, z["y"])


def test_device_fill_matches_host(gpu):
    L, r, offs, w, dense, x = setup("box2d1r", [133, 257], 5)
    with Plan(L, r, dense) as p:
        p.fill_random(5)
        assert np.array_equal(p.download(), fill_random_host(L, 5))
    assert np.array_equal(fill_random_host(L, 5), x)


def test_graph_equals_stream_and_is_deterministic(gpu):
    L, r, offs, w, dense, x = setup("box2d1r", [300, 520], 8)
    res = []
    for graph in (False, True, True):
        with Plan(L, r, dense, use_graph=graph) as p:
            p.upload(x)
            for _ in range(3):  # replays of one captured graph
                p.run(9)
            res.append(p.download())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[1], res[2])
    check(L, res[0], oracle.run(L, offs, w, x, 27))


def test_fusion_depths_agree(gpu):
    L, r, offs, w, dense, x = setup("box2d1r", [200, 333], 9)
    want = oracle.run(L, offs, w, x, 12)
    for kf in (1, 2, 3, 4):
        with Plan(L, r, dense, fusion_depth=kf) as p:
            p.upload(x)
            rep = p.run(12)
            assert rep.fusion_depth == kf
            check(L, p.download(), want)


def _window_check(L, full_out, x, offs, w, steps, r, y0, x0, size):
    """Recompute a size x size window of the final state from its light cone."""
    ny, nx = L.n[1], L.n[2]
    cone = steps * r
    wy0, wx0 = max(0, y0 - cone), max(0, x0 - cone)
    wy1, wx1 = min(ny, y0 + size + cone), min(nx, x0 + size + cone)
    sub = Layout.with_halo([wy1 - wy0, wx1 - wx0], r)
    xs = np.zeros(sub.cells)
    sub.interior(xs)[0] = L.interior(x)[0][wy0:wy1, wx0:wx1]
    ys = oracle.run(sub, offs, w, xs, steps, threads=0)
    want = sub.interior(ys)[0][y0 - wy0:y0 - wy0 + size, x0 - wx0:x0 - wx0 + size]
    got = L.interior(full_out)[0][y0:y0 + size, x0:x0 + size]
    rel_max, _ = rel_errors(got, want)
    return rel_max


def test_full_size_box2d1r_windows(gpu):
    # BASELINE config: box2d1r, 10240 x 10240, halo 1, T = 64
    n, steps = 10240, 64
    L, r, offs, w, dense, _ = setup("box2d1r", [8, 8], 1234)  # weights only
    L = Layout.with_halo([n, n], 1)
    x = fill_random_host(L, 1234)
    with Plan(L, r, dense, use_graph=True) as p:
        p.fill_random(1234)
        p.run(steps)
        out = p.download()
    worst = 0.0
    for (y0, x0) in [(0, 0), (0, n - 96), (n - 96, 0), (n - 96, n - 96), (5000, 3333), (64 * 37 + 3, 64 * 90 - 5)]:
        worst = max(worst, _window_check(L, out, x, offs, w, steps, r, y0, x0, 96))
    assert worst <= TOL, worst


def _stream_ids():
    import test_kernel_emulation as em
    return [(v["id"], v["name"], v["R"], v["shape"], v["kf"], v["by"], v["bx"]) for v in em.VARIANTS_STREAM]


@pytest.mark.parametrize("vid,vname,R,shape,kf,by,bx", _stream_ids())
def test_every_stream_variant(gpu, vid, vname, R, shape, kf, by, bx):
    # the pipelined persistent kernels (one workgroup per CU walking its tiles with two
    # LDS buffers): 3 x 3 tiles (edge and interior tiles), and grids with several tiles
    # per workgroup so that buffers are reused across interior and edge tiles
    print(f"variant {vname}", flush=True)
    name = {(1, 0): "box2d1r", (1, 1): "star2d1r", (3, 0): "box2d3r", (3, 1): "star2d3r"}[(R, shape)]
    for sizes in ([3 * by - 3, 3 * bx - 5], [1100, 1301], [2011, 2100]):
        L, R, offs, w, dense, x = setup(name, sizes, vid)
        for steps in (kf, 2 * kf + 1):
            with Plan(L, R, dense, variant=vid) as p:
                p.upload(x)
                p.run(steps)
                assert f"variant={vid}" in p.describe()
                got = p.download()
            check(L, got, oracle.run(L, offs, w, x, steps, threads=0))


def _reg_ids():
    import test_kernel_emulation as em
    return [(v["id"], v["name"], v["R"], v["shape"], v["kf"], v["bx"]) for v in em.VARIANTS_REG2D]


@pytest.mark.parametrize("vid,vname,R,shape,kf,ox", _reg_ids())
def test_every_reg_variant(gpu, vid, vname, R, shape, kf, ox):
    # the register-streaming kernels: grids of one to many strips (edge strips, the plain
    # path in tall chunks), a padded layout with odd halo widths, fused launches and remainders
    print(f"variant {vname}", flush=True)
    name = {(1, 0): "box2d1r", (1, 1): "star2d1r", (3, 0): "box2d3r", (3, 1): "star2d3r"}[(R, shape)]
    for sizes in ([3 * R * kf + 5, ox // 2 + 3], [1100, 1301], [2011, 2 * ox + 7], [7, 5000]):
        L, R, offs, w, dense, x = setup(name, sizes, vid)
        for steps in (kf, 2 * kf + 1):
            with Plan(L, R, dense, variant=vid) as p:
                p.upload(x)
                p.run(steps)
                assert f"variant={vid}" in p.describe()
                got = p.download()
            check(L, got, oracle.run(L, offs, w, x, steps, threads=0))
    L = Layout(2, [909, 1511], [4, 5], [899, 1500])
    ndim, R, _ = SHAPES[name]
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), vid + 1)
    x = oracle.fill_data(L, vid + 1)
    with Plan(L, R, dense_weights(ndim, R, offs, w), variant=vid) as p:
        p.upload(x)
        p.run(kf + 2)
        got = p.download()
    check(L, got, oracle.run(L, offs, w, x, kf + 2, threads=0))


@pytest.mark.parametrize("vid", [0, 811106, 750212, 700408, 711106])
def test_full_size_box2d1r_windows_variants(gpu, vid):
    # the measured configuration (box2d1r 10240^2, halo 1, T = 64) with the pipelined
    # kernel forced (0: the plan's default choice)
    n, steps = 10240, 64
    L, r, offs, w, dense, _ = setup("box2d1r", [8, 8], 1234)  # weights only
    L = Layout.with_halo([n, n], 1)
    x = fill_random_host(L, 1234)
    with Plan(L, r, dense, use_graph=True, variant=vid) as p:
        p.fill_random(1234)
        p.run(steps)
        out = p.download()
    worst = 0.0
    for (y0, x0) in [(0, 0), (0, n - 96), (n - 96, 0), (n - 96, n - 96), (5000, 3333), (64 * 37 + 3, 64 * 90 - 5)]:
        worst = max(worst, _window_check(L, out, x, offs, w, steps, r, y0, x0, 96))
    assert worst <= TOL, worst


@pytest.mark.parametrize("name", ["star2d3r", "box2d3r"])
def test_full_size_radius3_windows(gpu, name):
    # the radius-3 configuration: 8192^2, halo 3, T = 64, FP64 (autotuned plan, as measured),
    # checked on windows at the corners, edges and inside whose 192-cell light cone the oracle
    # recomputes exactly
    n, steps = 8192, 64
    L, r, offs, w, dense, _ = setup(name, [8, 8], 4321)
    L = Layout.with_halo([n, n], 3)
    x = fill_random_host(L, 4321)
    with Plan(L, r, dense, use_graph=True, autotune=1) as p:
        p.fill_random(4321)
        p.run(steps)
        out = p.download()
    worst = 0.0
    for (y0, x0) in [(0, 0), (n - 64, n - 64), (0, 4000), (4100, 2050)]:
        worst = max(worst, _window_check(L, out, x, offs, w, steps, r, y0, x0, 64))
    print(f"{name} 8192^2 T=64 worst window rel max {worst:.2e}", flush=True)
    assert worst <= TOL, worst
