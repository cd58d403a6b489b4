"""Pins the CPU oracle (oracle/cs_oracle.c) -- CPU only.

Checks, each chosen so that a plausible mistake (dropped tap, wrong sign or
index, transposed axes, halo overwritten, wrong ping-pong parity) fails one:
  * golden vectors written by an independent numpy restatement (bitwise);
  * brute force scalar loop at sampled points (bitwise);
  * impulse response == repeated self-convolution of the weights;
  * constant field stays constant farther than t*r from the boundary;
  * separable box weights == successive 1-D numpy.convolve passes;
  * the reference's 1d1r / 1d2r weight expansions == numpy.convolve of the taps;
  * SplitMix64 stream values against a pure-Python restatement.
"""
import glob
import os

import numpy as np
import pytest

from convstencil_amd.shapes import Layout, SHAPES, shape_offsets, offsets, expand_1d1r, expand_1d2r, compose
from oracle import oracle
from refimpl import run_numpy, brute_point, splitmix64_py

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _layout_from(z):
    return Layout(3, [int(v) for v in z["ext"]], [int(v) for v in z["org"]], [int(v) for v in z["n"]])


@pytest.mark.parametrize("path", sorted(p for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                                        if not p.endswith("splitmix64.npz")))
def test_golden_bitwise(path):
    z = np.load(path)
    L = _layout_from(z)
    offs = [tuple(int(v) for v in o) for o in z["offs"]]
    y = oracle.run(L, offs, z["w"], z["x"], int(z["steps"]))
    assert np.array_equal(y, z["y"]), os.path.basename(path)


def test_golden_inputs_match_generator():
    # the fixture inputs were written by the pure-Python SplitMix64; the oracle's
    # generator must produce the same bits
    for name in ["box2d1r", "star3d1r", "1d1r"]:
        z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
        L = _layout_from(z)
        assert np.array_equal(oracle.fill_data(L, int(z["seed"])), z["x"])
        assert np.array_equal(oracle.fill_weights(len(z["w"]), int(z["seed"])), z["w"])


def test_splitmix_stream_values():
    z = np.load(os.path.join(GOLDEN, "splitmix64.npz"))
    assert [int(v) for v in z["data"]] == oracle.splitmix_stream(1234, 0, 8)
    assert [int(v) for v in z["weights"]] == oracle.splitmix_stream(1234, 1, 8)
    # Steele et al.'s published first output for state 0 after one increment
    outs, _ = splitmix64_py(0, 1)
    assert outs[0] == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("name", list(SHAPES))
def test_brute_force_points(name):
    ndim, r, kind = SHAPES[name]
    sizes = {1: [29], 2: [9, 13], 3: [5, 6, 7]}[ndim]
    L = Layout.with_halo(sizes, r + 1)  # wider halo than needed: must stay untouched
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 5)
    x = oracle.fill_data(L, 5)
    # non-zero halo in buffer 0, zeros in buffer 1: the alternating halo matters
    x3 = x.reshape(L.ext).copy()
    rng = np.random.default_rng(0)
    mask = np.ones(L.ext, dtype=bool)
    o, n = L.org, L.n
    mask[o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]] = False
    x3[mask] = rng.uniform(-1, 1, size=int(mask.sum()))
    x = x3.reshape(-1)
    y1 = oracle.run(L, offs, w, x, 1)
    y2 = oracle.run(L, offs, w, x, 2)
    a3 = x.reshape(L.ext)
    y13 = y1.reshape(L.ext)
    y23 = y2.reshape(L.ext)
    # halo: after 1 step the state is buffer 1 (halo 0); after 2 steps buffer 0 (input halo)
    assert np.all(y13[mask] == 0.0)
    assert np.array_equal(y23[mask], a3[mask])
    for idx in [(0, 0, 0), tuple(v - 1 for v in n), tuple(v // 2 for v in n)]:
        z_, y_, x_ = (o[k] + idx[k] for k in range(3))
        assert y13[z_, y_, x_] == brute_point(a3, z_, y_, x_, offs, w)
        # step 2 reads buffer 1, whose halo is zero
        assert y23[z_, y_, x_] == brute_point(y13, z_, y_, x_, offs, w)
    assert np.array_equal(y2, run_numpy(x, L, offs, w, 2))


@pytest.mark.parametrize("name,steps", [("1d1r", 5), ("box2d1r", 3), ("star2d3r", 2), ("box3d1r", 2),
                                        ("star3d1r", 3)])
def test_impulse_response(name, steps):
    ndim, r, kind = SHAPES[name]
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 3)
    span = 2 * r * steps + 1
    sizes = [span + 4] * ndim
    L = Layout.with_halo(sizes, r)
    x = np.zeros(L.cells)
    c = [o + s // 2 for o, s in zip(L.org, L.n)]
    x.reshape(L.ext)[tuple(c)] = 1.0
    y = oracle.run(L, offs, w, x, steps).reshape(L.ext)
    d = 2 * r + 1
    cube = np.zeros((d,) * ndim)
    for (dz, dy, dx), v in zip(offs, w):
        cube[tuple([dz + r, dy + r, dx + r][3 - ndim:])] = v
    want = compose(cube, steps, ndim)
    # the response of u(p) = sum w(o) u(p+o) is the weight cube mirrored
    want = want[(slice(None, None, -1),) * ndim]
    h = r * steps
    sl = tuple(slice(ci - h, ci + h + 1) for ci in c[3 - ndim:])
    got = y[(0,) * (3 - ndim)][sl]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
    assert abs(y.sum() - 1.0) < 1e-13  # mass is conserved while the support stays inside


@pytest.mark.parametrize("name", ["box2d1r", "box2d3r", "star3d1r", "1d2r"])
def test_constant_field(name):
    ndim, r, kind = SHAPES[name]
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 9)
    steps = 3
    sizes = [2 * r * steps + 5] * ndim
    L = Layout.with_halo(sizes, r)
    x = np.zeros(L.cells)
    L.interior(x)[...] = 1.0
    y = oracle.run(L, offs, w, x, steps)
    inner = L.interior(y)
    h = r * steps
    core = inner[tuple(slice(h, -h) if L.n[k] > 1 else slice(None) for k in range(3))]
    np.testing.assert_allclose(core, 1.0, rtol=0, atol=4e-16 * steps * len(offs))
    # next to the zero halo the field must have decayed
    assert inner.reshape(-1)[0] < 1.0 - 1e-3


def test_separable_box_is_two_1d_passes():
    r = 1
    a = np.array([0.2, 0.5, 0.3])
    b = np.array([0.1, 0.7, 0.2])
    offs = offsets(2, r, "box")
    w = [a[dy + r] * b[dx + r] for (_, dy, dx) in offs]
    L = Layout.with_halo([12, 15], r)
    x = oracle.fill_data(L, 4)
    y = L.interior(oracle.run(L, offs, w, x, 1))[0]
    u = L.interior(x)[0]
    up = np.pad(u, 1)
    rows = np.array([np.convolve(row, b[::-1], mode="valid") for row in up])        # along x
    cols = np.array([np.convolve(col, a[::-1], mode="valid") for col in rows.T]).T  # along y
    np.testing.assert_allclose(y, cols, rtol=0, atol=1e-15)


def test_reference_expansions_are_convolutions():
    rng = np.random.default_rng(1)
    for _ in range(20):
        v = rng.uniform(-1, 1, 3)
        np.testing.assert_allclose(expand_1d1r(v), np.convolve(np.convolve(v, v), v), rtol=0, atol=1e-15)
        v5 = rng.uniform(-1, 1, 5)
        np.testing.assert_allclose(expand_1d2r(v5), np.convolve(v5, v5), rtol=0, atol=1e-15)
    # integer taps: exact
    v = np.array([1.0, 2.0, 3.0])
    assert np.array_equal(expand_1d1r(v), np.convolve(np.convolve(v, v), v))


def test_threads_do_not_change_bits():
    L = Layout.with_halo([40, 50], 3)
    offs = shape_offsets("box2d3r")
    w = oracle.fill_weights(len(offs), 2)
    x = oracle.fill_data(L, 2)
    assert np.array_equal(oracle.run(L, offs, w, x, 3, threads=1), oracle.run(L, offs, w, x, 3, threads=4))
