"""Autotuned plans on the GPU: a timed choice, its cache entry, the cached reuse, parity."""
import numpy as np
import pytest

from convstencil_amd import Plan, Layout, shape_offsets, dense_weights, SHAPES, _lib
from oracle import oracle
from refimpl import rel_errors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,prec,tol", [("box2d1r", "fp64", 1e-12), ("star2d3r", "fp64", 1e-12),
                                           ("box2d3r", "bf16x3", 1e-6), ("star3d1r", "fp64", 1e-12)])
def test_autotune_then_cached(gpu, tmp_path, monkeypatch, name, prec, tol):
    print(f"case autotune {name} {prec}", flush=True)
    monkeypatch.setenv("CONVSTENCIL_TUNE_CACHE", str(tmp_path / "tune.txt"))
    lib = _lib.load()
    lib.cs_tune_cache_reset()
    ndim, r, kind = SHAPES[name]
    L = Layout.with_halo([300, 257] if ndim == 2 else [40, 50, 70], r)
    offs = shape_offsets(name)
    w = oracle.fill_weights(len(offs), 9)
    dense = dense_weights(ndim, r, offs, w)
    x = oracle.fill_data(L, 9)
    want = oracle.run(L, offs, w, x, 10)
    variants = []
    for how in ("timed", "cached"):
        with Plan(L, r, dense, precision=prec, autotune=1) as p:
            d = p.describe()
            assert f"autotune={how}" in d, d
            variants.append(d.split()[0])
            p.upload(x)
            p.run(10)
            got = p.download()
        rm, rl2 = rel_errors(L.interior(got), L.interior(want))
        assert (rm <= tol) if prec == "fp64" else (rl2 <= tol), (how, rm, rl2)
    assert variants[0] == variants[1]
    assert (tmp_path / "tune.txt").read_text().count("\n") == 1
    monkeypatch.delenv("CONVSTENCIL_TUNE_CACHE")
    lib.cs_tune_cache_reset()
