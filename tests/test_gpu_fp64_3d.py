"""Plain FP64 3-D path on the GPU against the CPU oracle (<= 1e-12 relative max).

Covers star3d1r (7-point) and box3d1r (27-point) on ragged sizes and several
steps, the BASELINE size (512^3, T = 32) checked on windows whose light cone the
oracle recomputes exactly, and the reference's own launchers gpu_box_3d1r (343
points, 196 z-symmetric weights) and gpu_star_3d1r (63-point octahedron, 44
weights) with the reference's padded layout and a random non-zero halo.
"""
import numpy as np
import pytest

from convstencil_amd import Plan, Layout, shape_offsets, dense_weights, SHAPES, fill_random_host
from convstencil_amd import api
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check(L, got, want):
    rm, _ = rel_errors(L.interior(got), L.interior(want))
    assert rm <= TOL, rm


@pytest.mark.parametrize("name", ["star3d1r", "box3d1r"])
@pytest.mark.parametrize("sizes", [[7, 9, 11], [33, 40, 70], [3, 64, 128], [40, 33, 200]])
@pytest.mark.parametrize("steps", [1, 2, 5])
def test_parity(gpu, name, sizes, steps):
    print(f"case {name} {sizes} T={steps}", flush=True)
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo(sizes, r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), sum(sizes) + steps)
    x = oracle.fill_data(L, sum(sizes) + steps)
    with Plan(L, r, dense_weights(3, r, offs, w)) as p:
        p.upload(x)
        p.run(steps)
        got = p.download()
    _check(L, got, oracle.run(L, offs, w, x, steps))


@pytest.mark.parametrize("name", ["star3d1r", "box3d1r"])
def test_baseline_size_windows(gpu, name):
    n, steps = 512, 32
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo([n, n, n], r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 77)
    with Plan(L, r, dense_weights(3, r, offs, w), use_graph=True) as p:
        p.fill_random(77)
        p.run(steps)
        out = p.download()
    x = fill_random_host(L, 77)
    cone = steps * r
    worst = 0.0
    for (z0, y0, x0) in [(0, 0, 0), (n - 16, n - 16, n - 16), (200, 37, 451), (511 - 20, 100, 0)]:
        s = 16
        lo = [max(0, c - cone) for c in (z0, y0, x0)]
        hi = [min(n, c + s + cone) for c in (z0, y0, x0)]
        sub = Layout.with_halo([h - l for l, h in zip(lo, hi)], r)
        xs = np.zeros(sub.cells)
        sub.interior(xs)[...] = L.interior(x)[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        ys = sub.interior(oracle.run(sub, offs, w, xs, steps, threads=0))
        want = ys[z0 - lo[0]:z0 - lo[0] + s, y0 - lo[1]:y0 - lo[1] + s, x0 - lo[2]:x0 - lo[2] + s]
        got = L.interior(out)[z0:z0 + s, y0:y0 + s, x0:x0 + s]
        worst = max(worst, rel_errors(got, want)[0])
    assert worst <= TOL, worst


def _ref3d(h, m, n, seed):
    L = Layout(3, [h + 6, m + 6, n + 6], [3, 3, 3], [h, m, n])
    x = oracle.fill_data(L, seed)
    rng = np.random.default_rng(seed)
    a = x.reshape(L.ext)
    mask = np.ones(L.ext, dtype=bool)
    mask[3:3 + h, 3:3 + m, 3:3 + n] = False
    a[mask] = rng.integers(0, 100, int(mask.sum())).astype(np.float64)
    return L, a.reshape(-1)


def test_reference_gpu_box_3d1r(gpu):
    h, m, n, times = 9, 20, 70, 3
    L, x = _ref3d(h, m, n, 3)
    params = np.arange(196) / 100.0 / 50.0  # the reference's i/100 defaults, scaled to keep values small
    out = np.zeros(L.cells)
    api.gpu_box_3d1r(x, out, params, times, h, m, n)
    offs, w = [], []
    for dz in range(-3, 4):
        k = 3 - abs(dz)
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                offs.append((dz, dy, dx))
                w.append(params[k * 49 + (dy + 3) * 7 + (dx + 3)])
    want = oracle.run(L, offs, w, x, times)
    _check(L, out, want)
    m3 = np.ones(L.ext, dtype=bool)
    m3[3:3 + h, 3:3 + m, 3:3 + n] = False
    assert np.array_equal(out.reshape(L.ext)[m3], want.reshape(L.ext)[m3])


def test_reference_gpu_star_3d1r(gpu):
    h, m, n, times = 8, 17, 66, 2
    L, x = _ref3d(h, m, n, 4)
    params = np.linspace(0.001, 0.03, 44)
    out = np.zeros(L.cells)
    api.gpu_star_3d1r(x, out, params, times, h, m, n)
    base = {3: 0, 2: 1, 1: 6, 0: 19}
    offs, w = [], []
    for dz in range(-3, 4):
        r = 3 - abs(dz)
        idx = base[abs(dz)]
        for dy in range(-r, r + 1):
            span = r - abs(dy)
            for dx in range(-span, span + 1):
                offs.append((dz, dy, dx))
                w.append(params[idx])
                idx += 1
    want = oracle.run(L, offs, w, x, times)
    _check(L, out, want)


def _variants_3d():
    import ctypes
    from convstencil_amd import _lib
    lib = _lib.load()
    lib.cs_debug_variant_count.restype = ctypes.c_int32
    lib.cs_debug_variant_info.restype = ctypes.c_int32
    lib.cs_debug_variant_info.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_char_p, ctypes.c_int32]
    out = []
    for i in range(lib.cs_debug_variant_count(3)):
        info = (ctypes.c_int64 * 12)()
        name = ctypes.create_string_buffer(128)
        lib.cs_debug_variant_info(3, i, info, name, 128)
        out.append((int(info[0]), int(info[2]), int(info[3]), name.value.decode()))
    return out


# radius-1 box and star variants of both families (the radius-1 diamond is the star)
VARIANTS_3D = [v for v in _variants_3d() if v[1] == 1 and v[2] in (0, 1)]


@pytest.mark.parametrize("vid,R,shape,vname", VARIANTS_3D, ids=[v[3] for v in VARIANTS_3D])
def test_every_variant(gpu, vid, R, shape, vname):
    print(f"case {vname}", flush=True)
    name = "box3d1r" if shape == 0 else "star3d1r"
    L = Layout.with_halo([37, 45, 140], 1)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), vid)
    x = oracle.fill_data(L, vid)
    with Plan(L, 1, dense_weights(3, 1, offs, w), variant=vid) as p:
        assert f"variant={vid}" in p.describe()
        p.upload(x)
        p.run(5)
        got = p.download()
    _check(L, got, oracle.run(L, offs, w, x, 5))


def _reg_variants_3d(fused=False):
    import ctypes
    from convstencil_amd import _lib
    lib = _lib.load()
    lib.cs_debug_variant_count.restype = ctypes.c_int32
    lib.cs_debug_variant_info.restype = ctypes.c_int32
    lib.cs_debug_variant_info.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_char_p, ctypes.c_int32]
    out = []
    for i in range(lib.cs_debug_variant_count(9)):
        info = (ctypes.c_int64 * 12)()
        name = ctypes.create_string_buffer(128)
        lib.cs_debug_variant_info(9, i, info, name, 128)
        if (int(info[4]) > 1) == fused:
            out.append((int(info[0]), int(info[3]), name.value.decode()))
    return out


REG_3D = _reg_variants_3d()
FUSED_3D = _reg_variants_3d(True)


# the register-streaming single-step kernels (k3d_reg.hip, Variant::mf 9): plain-path
# waves inside the large grid, ragged strips and tiles, halos of odd widths
@pytest.mark.parametrize("vid,shape,vname", REG_3D, ids=[v[2] for v in REG_3D])
def test_every_reg_variant(gpu, vid, shape, vname):
    name = "box3d1r" if shape == 0 else "star3d1r"
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), vid)
    for L, steps in [(Layout.with_halo([50, 21, 600], 1), 1), (Layout.with_halo([37, 45, 140], 1), 5),
                     (Layout(3, [29, 14, 300], [2, 1, 3], [25, 11, 290]), 3), (Layout.with_halo([3, 1, 5], 1), 2)]:
        print(f"case {vname} {L.n} T={steps}", flush=True)
        x = oracle.fill_data(L, vid + steps)
        with Plan(L, 1, dense_weights(3, 1, offs, w), variant=vid) as p:
            assert f"variant={vid}" in p.describe()
            p.upload(x)
            p.run(steps)
            got = p.download()
        _check(L, got, oracle.run(L, offs, w, x, steps))


# the fused register-streaming kernels (two or three steps per launch; plans fuse with a zero halo,
# which the seeded fill leaves): plain-path waves in the large grid, ragged strips and
# tiles, halos of odd widths, even and odd step counts (a single-step remainder)
@pytest.mark.parametrize("vid,shape,vname", FUSED_3D, ids=[v[2] for v in FUSED_3D])
def test_every_fused_reg_variant(gpu, vid, shape, vname):
    name = "box3d1r" if shape == 0 else "star3d1r"
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), vid)
    for L, steps in [(Layout.with_halo([60, 30, 600], 1), 4), (Layout.with_halo([37, 45, 140], 1), 5),
                     (Layout(3, [29, 14, 300], [2, 1, 3], [25, 11, 290]), 2), (Layout.with_halo([3, 1, 5], 1), 2),
                     (Layout.with_halo([1, 2, 130], 2), 3), (Layout.with_halo([7, 9, 11], 1), 5)]:
        print(f"case {vname} {L.n} T={steps}", flush=True)
        x = oracle.fill_data(L, vid + steps)
        with Plan(L, 1, dense_weights(3, 1, offs, w), variant=vid) as p:
            assert f"variant={vid}" in p.describe() and "halo_uniform=1" in p.describe()
            p.upload(x)
            p.run(steps)
            got = p.download()
        _check(L, got, oracle.run(L, offs, w, x, steps))


@pytest.mark.parametrize("name", ["star3d1r", "box3d1r"])
def test_nonzero_halo_runs_single_steps(gpu, name):
    # a non-zero halo: the plan (whose main variant fuses) runs every step on its single-step
    # kernel, which reads the B fragments the plan appends after the fused kernel's weights;
    # the halo stays as uploaded in both buffers' reads, as in the reference's ping-pong
    print(f"case {name} nonzero halo", flush=True)
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo([9, 13, 150], r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 91)
    x = oracle.fill_data(L, 91)
    rng = np.random.default_rng(91)
    a = x.reshape(L.ext)
    m = np.ones(L.ext, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = False
    a[m] = rng.uniform(-1.0, 1.0, int(m.sum()))
    x = a.reshape(-1)
    with Plan(L, r, dense_weights(3, r, offs, w)) as p:
        assert "halo_uniform" in p.describe()
        p.upload(x)
        assert "halo_uniform=0" in p.describe()
        p.run(5)
        got = p.download()
    _check(L, got, oracle.run(L, offs, w, x, 5))


# the register-blocked FMA kernels (k3d_fma.hip, Variant::mf 3): radius 2 and 3, each
# shape class, partial tiles and z-chunks
FMA_3D = [(35302, 3, "box"), (35322, 3, "diamond"), (35312, 3, "star"), (35303, 3, "box"), (35323, 3, "diamond"),
          (35305, 3, "box"), (35325, 3, "diamond"), (35300, 3, "box"), (35310, 3, "star"), (35320, 3, "diamond"),
          (35301, 3, "box"), (35321, 3, "diamond"), (35202, 2, "box"), (35212, 2, "star"), (35222, 2, "diamond"),
          (35200, 2, "box"), (35210, 2, "star"), (35220, 2, "diamond"), (35306, 3, "box")]


def _offsets3(r, kind):
    rr = range(-r, r + 1)
    keep = {"box": lambda a, b, c: True,
            "star": lambda a, b, c: (a != 0) + (b != 0) + (c != 0) <= 1,
            "diamond": lambda a, b, c: abs(a) + abs(b) + abs(c) <= r}[kind]
    return [(a, b, c) for a in rr for b in rr for c in rr if keep(a, b, c)]


@pytest.mark.parametrize("vid,r,kind", FMA_3D)
def test_fma_variants(gpu, vid, r, kind):
    print(f"case k3d_fma {vid} r{r} {kind}", flush=True)
    offs = _offsets3(r, kind)
    for sizes in ([11, 13, 70], [40, 33, 65], [3, 70, 129]):
        L = Layout.with_halo(sizes, r)
        w = oracle.fill_weights(len(offs), vid + sizes[0])
        x = oracle.fill_data(L, vid + sizes[1])
        with Plan(L, r, dense_weights(3, r, offs, w), variant=vid) as p:
            assert f"variant={vid}" in p.describe()
            p.upload(x)
            p.run(3)
            got = p.download()
        _check(L, got, oracle.run(L, offs, w, x, 3))


def test_fma_is_the_radius3_default(gpu):
    # radius 3 has no 8 x 8 block kernel: plans take the FMA family, box and octahedron alike
    for kind, vid in (("box", 35302), ("diamond", 35322), ("star", 35312)):
        offs = _offsets3(3, kind)
        L = Layout.with_halo([8, 20, 70], 3)
        with Plan(L, 3, dense_weights(3, 3, offs, oracle.fill_weights(len(offs), 5))) as p:
            assert f"variant={vid}" in p.describe(), p.describe()


@pytest.mark.parametrize("ndim,sizes", [(1, [0]), (2, [0, 7]), (2, [5, 0]), (3, [3, 0, 4]), (3, [0, 2, 2])])
@pytest.mark.parametrize("steps", [1, 2, 3])
def test_empty_interior(gpu, ndim, sizes, steps):
    # no interior cell: the steps only swap the ping-pong buffers, as the oracle's loop does
    name = {1: "1d1r", 2: "box2d1r", 3: "box3d1r"}[ndim]
    L = Layout.with_halo(sizes, 1)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 9)
    x = np.random.default_rng(steps).uniform(-1, 1, L.cells)  # a halo that is not zero
    with Plan(L, 1, dense_weights(ndim, 1, offs, w)) as p:
        p.upload(x)
        p.run(steps)
        got = p.download()
    want = oracle.run(L, offs, w, x, steps)  # odd steps: the zeroed second buffer
    assert got.tobytes() == want.tobytes()
