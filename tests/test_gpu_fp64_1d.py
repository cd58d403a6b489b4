"""Plain FP64 1-D path on the GPU against the CPU oracle (<= 1e-12 relative max).

Covers 1d1r (3-point) and 1d2r (5-point) physical stencils at every fusion
depth and remainder, ragged sizes, the BASELINE size (2^20 points, T = 64) in
full, and the reference's own launchers gpu_1d1r / gpu_star_1d2r_step2 with
their padded layouts, random non-zero halos and pre-fused 7- / 9-tap weights.
"""
import numpy as np
import pytest

from convstencil_amd import Plan, Layout, shape_offsets, dense_weights, SHAPES, fill_random_host
from convstencil_amd import api
from convstencil_amd.shapes import expand_1d1r, expand_1d2r
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _check(L, got, want):
    rm, _ = rel_errors(L.interior(got), L.interior(want))
    assert rm <= TOL, rm


@pytest.mark.parametrize("name", ["1d1r", "1d2r"])
@pytest.mark.parametrize("n", [1, 63, 2048, 4097, 100003])
@pytest.mark.parametrize("steps", [1, 3, 33, 64])
def test_parity(gpu, name, n, steps):
    print(f"case {name} n={n} T={steps}", flush=True)
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo([n], r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), n + steps)
    x = oracle.fill_data(L, n + steps)
    with Plan(L, r, dense_weights(1, r, offs, w)) as p:
        p.upload(x)
        p.run(steps)
        got = p.download()
    _check(L, got, oracle.run(L, offs, w, x, steps))


def test_baseline_size_full(gpu):
    n, steps = 1 << 20, 64
    L = Layout.with_halo([n], 1)
    offs = shape_offsets("1d1r")
    w = oracle.fill_weights(3, 2024)
    with Plan(L, 1, dense_weights(1, 1, offs, w), use_graph=True) as p:
        p.fill_random(2024)
        rep = p.run(steps)
        got = p.download()
    x = fill_random_host(L, 2024)
    _check(L, got, oracle.run(L, offs, w, x, steps, threads=0))
    assert rep.launches <= 4


def _rand_halo(L, x, seed):
    rng = np.random.default_rng(seed)
    a = x.reshape(L.ext).copy()
    m = np.ones(L.ext, dtype=bool)
    m[0, 0, L.org[2]:L.org[2] + L.n[2]] = False
    a[m] = rng.integers(0, 10000, int(m.sum())).astype(np.float64)
    return a.reshape(-1)


def test_reference_gpu_1d1r(gpu):
    # host arrays of n + 7, interior at [4, n + 4), output shifted by one (src/1d/gpu_1r.cu:266,299)
    n, times = 3000, 4
    params = expand_1d1r([0.25, 0.5, 0.25])
    L = Layout(1, [n + 7], [4], [n])
    x = _rand_halo(L, oracle.fill_data(L, 5), 5)
    out = np.zeros(n + 7)
    api.gpu_1d1r(x, out, params, times, n)
    offs = [(0, 0, d) for d in range(-3, 4)]
    want = oracle.run(L, offs, params, x, times)
    got = np.empty(n + 7)
    got[1:] = out[:n + 6]
    got[0] = want[0]
    _check(L, got, want)
    assert np.array_equal(out[:3], want[1:4]) and np.array_equal(out[n + 3:n + 6], want[n + 4:n + 7])


def test_reference_gpu_1d2r(gpu):
    # host arrays of n + 8, interior [4, n + 4), output = first n + 7 elements (src/1d/gpu_2r.cu:115,156)
    n, times = 2500, 3
    params = expand_1d2r([0.1, 0.2, 0.4, 0.2, 0.1])
    L = Layout(1, [n + 8], [4], [n])
    x = _rand_halo(L, oracle.fill_data(L, 6), 6)
    out = np.zeros(n + 8)
    api.gpu_star_1d2r_step2(x, out, params, times, n)
    offs = [(0, 0, d) for d in range(-4, 5)]
    want = oracle.run(L, offs, params, x, times)
    got = want.copy()
    got[:n + 7] = out[:n + 7]
    _check(L, got, want)
    assert np.array_equal(out[:4], want[:4])


# 6xxxxx: the register-resident kernels, 600000 + R * 10000 + cells per lane * 100 + depth
@pytest.mark.parametrize("vid", [511016, 521016, 531016, 511008, 511032, 531008, 933001, 611864, 611032, 612664, 611064, 610864,
                                 611808, 621832, 621808, 631816, 631804, 641816, 641802])
def test_tile_variants(gpu, vid):
    # every alternative tile, forced: partial tiles, fused launches and remainders
    print(f"case variant {vid}", flush=True)
    r = 3 if vid == 933001 else (vid // 10000) % 10 if vid >= 600000 else 1
    offs = [(0, 0, d) for d in range(-r, r + 1)]
    for n, steps in ((5000, 37), (513, 16), (100003, 9)):
        L = Layout.with_halo([n], r)
        w = oracle.fill_weights(len(offs), vid + n)
        x = oracle.fill_data(L, vid + steps)
        with Plan(L, r, dense_weights(1, r, offs, w), variant=vid) as p:
            assert f"variant={vid}" in p.describe()
            p.upload(x)
            p.run(steps)
            got = p.download()
        _check(L, got, oracle.run(L, offs, w, x, steps))
