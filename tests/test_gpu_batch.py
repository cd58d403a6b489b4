"""Batched persistent kernel (cs_batch_*) on the GPU, every grid against the CPU oracle.

One persistent launch advances many independent 2-D grids of uneven sizes and
mixed stencils (box2d1r, star2d3r); each grid must match what the single-grid
path and the oracle give for it alone: FP64 to 1e-12 relative (max), BF16x3 to
the split-precision bar (relative L2 <= 1e-6, relative max <= 5e-6).
"""
import numpy as np
import pytest

from convstencil_amd import Batch, Layout, shape_offsets, dense_weights, SHAPES
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu

# (stencil, interior sizes): partial tiles, a grid smaller than one tile, one exact tile,
# and 3 x 3 tiles whose centre tile lies inside the interior (the unmasked tile path)
GRIDS = [("box2d1r", [64, 64]), ("star2d3r", [100, 130]), ("box2d1r", [257, 75]), ("star2d3r", [31, 200]),
         ("box2d1r", [3, 5]), ("star2d3r", [150, 64]), ("box2d1r", [130, 190]), ("box2d1r", [190, 203]),
         ("star2d3r", [189, 200])]


def _grid(name, sizes, seed, random_halo=False):
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo(sizes, r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), seed)
    x = oracle.fill_data(L, seed)
    if random_halo:
        rng = np.random.default_rng(seed)
        m = np.ones(L.ext, dtype=bool)
        m[0, r:r + sizes[0], r:r + sizes[1]] = False
        x = x.reshape(L.ext)
        x[m] = rng.uniform(-1, 1, int(m.sum()))
        x = x.reshape(-1)
    return L, r, offs, w, dense_weights(2, r, offs, w), x


def _check(batch, grids, steps, prec):
    for i, (L, r, offs, w, dense, x) in enumerate(grids):
        got = batch.download(i)
        want = oracle.run(L, offs, w, x, steps)
        rm, rl2 = rel_errors(L.interior(got), L.interior(want))
        if prec == "fp64":
            assert rm <= 1e-12, (i, rm)
        else:
            assert rl2 <= 1e-6 and rm <= 5e-6, (i, rm, rl2)
        # halo cells are those of the buffer the state ends in, exactly
        assert np.array_equal(got[~L.interior_mask()], want[~L.interior_mask()]), i


@pytest.mark.parametrize("prec", ["fp64", "bf16x3", "fp16x2"])
@pytest.mark.parametrize("steps", [1, 7, 12])
def test_batch_matches_oracle(gpu, prec, steps):
    print(f"case batch {prec} T={steps}", flush=True)
    grids = [_grid(n, s, 100 + k) for k, (n, s) in enumerate(GRIDS)]
    with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids], precision=prec) as b:
        for i, g in enumerate(grids):
            b.upload(i, g[5])
        b.run(steps)
        _check(b, grids, steps, prec)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_consecutive_runs(gpu, graph):
    # the state carries over between runs (buffer parity per grid); 5 + 3 + 5 steps
    print(f"case batch consecutive graph={graph}", flush=True)
    grids = [_grid(n, s, 200 + k) for k, (n, s) in enumerate(GRIDS)]
    with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids], use_graph=graph) as b:
        for i, g in enumerate(grids):
            b.upload(i, g[5])
        b.run(5)
        b.run(3)
        b.run(5)
        _check(b, grids, 13, "fp64")


def test_batch_nonzero_halo(gpu):
    # a grid with a non-zero halo runs single steps (the halo alternates with the zero
    # halo of the second buffer); the other grids still fuse
    print("case batch nonzero halo", flush=True)
    grids = [_grid(n, s, 300 + k, random_halo=(k == 1)) for k, (n, s) in enumerate(GRIDS[:4])]
    with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids]) as b:
        for i, g in enumerate(grids):
            b.upload(i, g[5])
        b.run(6)
        _check(b, grids, 6, "fp64")


def test_batch_device_fill_and_time(gpu):
    # device-side fill equals the host generator; timing runs and reports
    print("case batch fill/time", flush=True)
    grids = [_grid(n, s, 400 + k) for k, (n, s) in enumerate(GRIDS[:3])]
    with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids], use_graph=True) as b:
        b.fill_random([7, 8, 9])
        for i, g in enumerate(grids):
            assert np.array_equal(b.download(i), oracle.fill_data(g[0], [7, 8, 9][i]))
        rep = b.time(8, warmup=1, reps=2)
        assert rep.elapsed_ms > 0 and rep.gstencil_per_s > 0
        assert "grids=3" in b.describe()


def test_batch_rejects_unsupported(gpu):
    # a radius-2 box has no batch tile body: creation fails with a status, no launch
    L = Layout.with_halo([40, 40], 2)
    dense = np.full(25, 1.0 / 25)
    with pytest.raises(RuntimeError):
        Batch([L], [2], [dense])


def test_batch_rejects_split_forms_without_bodies(gpu):
    # FP16 with three weight slices and FP32x2 have no batch tile bodies
    L = Layout.with_halo([40, 40], 1)
    dense = np.full(9, 1.0 / 9)
    for prec in ("fp16w3", "fp32x2"):
        with pytest.raises(RuntimeError):
            Batch([L], [1], [dense], precision=prec)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x2"])
def test_batch_split_error_after_100_steps(gpu, prec):
    """The batched workload's own depth: 24 grids, sides log-uniform in [128, 512], stencils
    drawn from {box2d1r, star2d3r} with their own seeded weights, T = 100 -- every grid
    within the split bar against the oracle.  Over 100 steps a weight error that repeats
    every step (its mass and first moments) and slices that carry the sign of the state
    (with the level-3 products dropped) would both grow with T."""
    rng = np.random.default_rng(77)
    G, T = 24, 100
    sides = np.exp(rng.uniform(np.log(128), np.log(512), (G, 2))).astype(int)
    names = rng.choice(["box2d1r", "star2d3r"], G)
    grids = [_grid(str(names[i]), [int(sides[i, 0]), int(sides[i, 1])], 500 + i) for i in range(G)]
    print(f"case batch {prec} T={T} grids={G}", flush=True)
    with Batch([g[0] for g in grids], [g[1] for g in grids], [g[4] for g in grids], precision=prec,
               use_graph=True) as b:
        for i, g in enumerate(grids):
            b.upload(i, g[5])
        b.run(T)
        got = [b.download(i) for i in range(G)]
    threads = oracle.max_threads()
    worst = (0.0, 0.0)
    for i, (L, r, offs, w, dense, x) in enumerate(grids):
        want = oracle.run(L, offs, w, x, T, threads=threads)
        rm, rl2 = rel_errors(L.interior(got[i]), L.interior(want))
        worst = (max(worst[0], rm), max(worst[1], rl2))
        assert rl2 <= 1e-6 and rm <= 5e-6, (i, names[i], sides[i], rm, rl2)
    print(f"worst rel max {worst[0]:.2e} rel L2 {worst[1]:.2e}", flush=True)
