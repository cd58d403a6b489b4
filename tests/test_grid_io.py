"""Grid files (cs_grid_save_txt / cs_grid_save_f64 / cs_grid_load_f64) -- CPU only,
except the command-line dumps at the end (GPU).

The text format is the reference's save_to_txt (src/1d/main.cu:21-36,
src/2d/main.cu:36-55, src/3d/main.cu:22-40): every cell of the padded array printed
with "%.0f" -- 1-D "index value" lines, 2-D tab-terminated rows, 3-D the rows of each
plane followed by a blank line and "Next".  The expected files here are built
independently with Python's "%.0f" formatting (round half to even, as C's printf).
"""
import os
import subprocess

import numpy as np
import pytest

import convstencil_amd as cs
from convstencil_amd import Layout, Plan, shape_offsets, dense_weights, SHAPES
from convstencil_amd._lib import ConvStencilError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "convstencil_amd", "bin")


def _expected_txt(layout, a):
    a = np.asarray(a).reshape(layout.ext)
    if layout.ndim == 1:
        return "".join(f"{i} {v:.0f}\n" for i, v in enumerate(a.reshape(-1)))
    out = []
    for z in range(layout.ext[0]):
        for y in range(layout.ext[1]):
            out.append("".join(f"{v:.0f}\t" for v in a[z, y]) + "\n")
        if layout.ndim == 3:
            out.append("\nNext\n")
    return "".join(out)


@pytest.mark.parametrize("sizes,halo", [([9], 3), ([4, 5], 1), ([3, 2, 4], 1), ([1, 1], 2)])
def test_text_dump_matches_reference_format(tmp_path, sizes, halo):
    L = Layout.with_halo(sizes, halo)
    rng = np.random.default_rng(len(sizes) * 10 + halo)
    a = rng.integers(-500, 500, L.cells).astype(np.float64) * 0.5  # halves: round-half-even cases
    path = tmp_path / "out.txt"
    cs.save_txt(path, L, a)
    assert path.read_text() == _expected_txt(L, a)


def test_text_dump_2d_literal(tmp_path):
    # a 2 x 3 padded array, written out by hand in the reference's format
    L = Layout(2, [2, 3], [0, 0], [2, 3])
    cs.save_txt(tmp_path / "o.txt", L, np.array([0.0, 1.5, 2.5, -3.0, 99.49, 1e3]))
    assert (tmp_path / "o.txt").read_text() == "0\t2\t2\t\n-3\t99\t1000\t\n"


@pytest.mark.parametrize("sizes,halo,extra", [([17], 1, 0), ([5, 7], 3, 2), ([3, 4, 5], 1, 0)])
def test_f64_round_trip_is_exact(tmp_path, sizes, halo, extra):
    L = Layout.with_halo(sizes, halo, extra_x=extra)
    rng = np.random.default_rng(3)
    a = rng.standard_normal(L.cells)
    a[:4] = [np.inf, -np.inf, -0.0, 5e-324]
    a[4] = np.nan
    cs.save_f64(tmp_path / "g.f64", L, a)
    L2, b = cs.load_f64(tmp_path / "g.f64")
    assert (L2.ndim, L2.ext, L2.org, L2.n) == (L.ndim, L.ext, L.org, L.n)
    assert a.tobytes() == b.tobytes()
    # header (96 bytes) + the padded array, nothing else
    assert os.path.getsize(tmp_path / "g.f64") == 96 + 8 * L.cells
    assert (tmp_path / "g.f64").read_bytes()[:8] == b"CSGRID64"


def test_f64_errors(tmp_path):
    L = Layout.with_halo([4, 4], 1)
    a = np.zeros(L.cells)
    with pytest.raises(ConvStencilError):
        cs.load_f64(tmp_path / "missing.f64")
    (tmp_path / "foreign.f64").write_bytes(b"NOTAGRID" + bytes(200))
    with pytest.raises(ConvStencilError):
        cs.load_f64(tmp_path / "foreign.f64")
    cs.save_f64(tmp_path / "t.f64", L, a)
    data = (tmp_path / "t.f64").read_bytes()
    (tmp_path / "trunc.f64").write_bytes(data[:-8])
    with pytest.raises(ConvStencilError):
        cs.load_f64(tmp_path / "trunc.f64")
    # an inconsistent layout (interior beyond the padded extent) is refused on save
    bad = Layout(2, [4, 4], [1, 1], [4, 4])
    with pytest.raises(ConvStencilError):
        cs.save_f64(tmp_path / "bad.f64", bad, np.zeros(bad.cells))
    with pytest.raises(ConvStencilError):
        cs.save_txt(tmp_path / "no_such_dir" / "x.txt", L, a)


def test_f64_capacity_checked(tmp_path):
    import ctypes
    from convstencil_amd import _lib
    L = Layout.with_halo([6], 1)
    cs.save_f64(tmp_path / "c.f64", L, np.arange(L.cells, dtype=np.float64))
    lib = _lib.load()
    cl = _lib.cs_layout()
    small = np.zeros(L.cells - 1)
    rc = lib.cs_grid_load_f64(os.fsencode(tmp_path / "c.f64"), ctypes.byref(cl),
                              small.ctypes.data_as(_lib.c_double_p), small.size)
    assert rc == _lib.CS_ERR_INVALID


@pytest.mark.parametrize("ext,org,n", [
    # 2^32 * 2^32 wraps a 64-bit product to 0; with ext[0] = 2 it would wrap to a small positive count
    ([2, 1 << 32, 1 << 32], [0, 0, 0], [2, 4, 4]),
    ([1, 1 << 31, (1 << 33) + 1], [0, 0, 0], [1, 2, 2]),
    # cell count fits but its byte count does not
    ([1, 1 << 31, 1 << 31], [0, 0, 0], [1, 2, 2]),
    # org + n overflows int64
    ([1, 4, 8], [0, 0, (1 << 63) - 1], [1, 2, 2]),
])
def test_f64_header_with_overflowing_extents_is_refused(tmp_path, ext, org, n):
    """A crafted header whose extents overflow the cell or byte count is refused before any
    read (a wrapped product could otherwise pass the capacity check)."""
    import ctypes
    import struct
    from convstencil_amd import _lib
    hdr = b"CSGRID64" + struct.pack("<II", 1, 3) + struct.pack("<9q", *ext, *org, *n) + bytes(8)
    assert len(hdr) == 96
    (tmp_path / "o.f64").write_bytes(hdr + bytes(8 * 16))
    lib = _lib.load()
    cl = _lib.cs_layout()
    buf = np.zeros(16)
    rc = lib.cs_grid_load_f64(os.fsencode(tmp_path / "o.f64"), ctypes.byref(cl),
                              buf.ctypes.data_as(_lib.c_double_p), buf.size)
    assert rc == _lib.CS_ERR_INVALID
    with pytest.raises(ConvStencilError):
        cs.load_f64(tmp_path / "o.f64")


# ---------------------------------------------------------------- command lines (GPU)

@pytest.mark.gpu
@pytest.mark.parametrize("prog,shape,sizes", [("convstencil_2d", "box2d1r", [40, 56]),
                                              ("convstencil_1d", "1d1r", [300]),
                                              ("convstencil_3d", "star3d1r", [6, 10, 12])])
def test_cli_dumps_match_plan(gpu, tmp_path, prog, shape, sizes):
    # physical mode, 5 steps from seed 7: the CLI's binary and text dumps equal the plan
    # API's result bit for bit
    ndim, r, kind = SHAPES[shape]
    txt, f64 = tmp_path / "o.txt", tmp_path / "o.f64"
    p = subprocess.run([os.path.join(BIN, prog), shape, *map(str, sizes), "5", "--steps=5", "--seed=7",
                        f"--write-output={txt}", f"--save-f64={f64}"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert f"Writing {txt}" in p.stdout
    L, got = cs.load_f64(f64)
    offs = shape_offsets(shape)
    dense = dense_weights(ndim, r, offs, cs.fill_weights_host(len(offs), 7))
    with Plan(Layout.with_halo(sizes, r), r, dense) as plan:
        plan.fill_random(7)
        plan.run(5)
        want = plan.download()
    assert (L.ext, L.org, L.n) == (Layout.with_halo(sizes, r).ext, Layout.with_halo(sizes, r).org,
                                   Layout.with_halo(sizes, r).n)
    assert got.tobytes() == want.tobytes()
    assert txt.read_text() == _expected_txt(L, want)


@pytest.mark.gpu
def test_cli_reference_mode_dump(gpu, tmp_path):
    # reference mode writes output_gpu.txt in the working directory, as the reference's
    # WRITE_OUTPUT build does: the padded (m + 6) x (n + 8) array
    p = subprocess.run([os.path.join(BIN, "convstencil_2d"), "box2d1r", "16", "24", "2", "--write-output"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0, p.stderr
    assert "Writing output_gpu.txt" in p.stdout
    rows = (tmp_path / "output_gpu.txt").read_text().split("\n")
    assert len(rows) == 16 + 6 + 1 and rows[-1] == ""
    assert all(len(r.split("\t")) == 24 + 8 + 1 for r in rows[:-1])
