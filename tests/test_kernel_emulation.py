"""CPU emulation of the 2-D FP64 MFMA kernels (k2d_f64.hip) -- no GPU needed.

For every compiled variant this replays the kernel's tile loop, LDS addressing,
block coverage, halo masking and MFMA lane algebra in numpy, with:
  * the k-step tables and B fragments taken from the library itself
    (include/convstencil_debug.h), so the tables that run on the GPU are checked;
  * the MFMA operand/result lane layouts measured on gfx950 (see the layout test
    in tests/test_gpu_mfma_layout.py):
      v_mfma_f64_16x16x4: A[l%16][l/16], B[l/16][l%16], D[4v + l/16][l%16]
      v_mfma_f64_4x4x4:   block (l>>2)&3, A[l&3][l>>4], B[l>>4][l&3], D[l>>4][l&3];
  * an assertion on every LDS and global index (in bounds of the buffers the
    kernel allocates), so an out-of-bounds access is found here, not on a GPU;
  * the result compared with the CPU oracle (1e-13 relative).
"""
import ctypes

import numpy as np
import pytest

from convstencil_amd import _lib
from convstencil_amd.shapes import Layout
from oracle import oracle


def _dbg():
    lib = _lib.load()
    lib.cs_debug_variant_count.restype = ctypes.c_int32
    lib.cs_debug_variant_info.restype = ctypes.c_int32
    lib.cs_debug_variant_info.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p,
                                          ctypes.c_int32]
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_ktable.restype = ctypes.c_int32
    lib.cs_debug_ktable.argtypes = [ctypes.c_int32] * 3 + [ip, ip, ip, ctypes.c_int32]
    lib.cs_debug_bfrag_2d.restype = ctypes.c_int32
    lib.cs_debug_bfrag_2d.argtypes = [ctypes.c_int32] * 3 + [_lib.c_double_p, _lib.c_double_p, ctypes.c_int32]
    return lib


def variants():
    lib = _dbg()
    out = []
    for i in range(lib.cs_debug_variant_count()):
        info = (ctypes.c_int64 * 10)()
        name = ctypes.create_string_buffer(128)
        assert lib.cs_debug_variant_info(i, info, name, 128) == 0
        keys = ["id", "mf", "R", "shape", "kf", "by", "bx", "nw", "lds", "px"]
        d = dict(zip(keys, [int(v) for v in info]))
        d["name"] = name.value.decode()
        out.append(d)
    return out


def ktable(mf, R, shape):
    lib = _dbg()
    cap = 256
    dy, c0, live = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
    n = lib.cs_debug_ktable(mf, R, shape, dy, c0, live, cap)
    assert n > 0
    return list(dy[:n]), list(c0[:n]), list(live[:n])


def bfrag(mf, R, shape, w):
    lib = _dbg()
    cap = 64 * 256
    out = np.zeros(cap)
    wd = np.ascontiguousarray(w, dtype=np.float64)
    nks = lib.cs_debug_bfrag_2d(mf, R, shape, wd.ctypes.data_as(_lib.c_double_p), out.ctypes.data_as(_lib.c_double_p),
                                cap)
    assert nks > 0
    return out[:nks * 64].reshape(nks, 64)


def lane_offsets(v, tab):
    """Per-lane LDS offsets off[ks][lane] relative to the block's (y, x) cell."""
    dy, c0, live = tab
    px = v["px"]
    lanes = np.arange(64)
    if v["mf"] == 16:
        return np.array([((lanes & 15) + dy[ks]) * px + c0[ks] + (lanes >> 4) for ks in range(len(dy))])
    nks = len(dy) // 4
    k, b4, i = lanes >> 4, (lanes >> 2) & 3, lanes & 3
    f = 4 * np.arange(nks)[:, None] + k[None, :]
    dyf = np.array(dy)[f]
    cf = np.array(c0)[f]
    return (i[None, :] + dyf) * px + 4 * b4[None, :] + cf


def mfma_block(v, a_lanes, bf):
    """a_lanes[ks][lane], bf[ks][lane] -> output value per (lane, v) as dict of cell offsets."""
    lanes = np.arange(64)
    if v["mf"] == 16:
        acc = np.zeros((16, 16))
        for ks in range(a_lanes.shape[0]):
            A = np.zeros((16, 4))
            B = np.zeros((4, 16))
            A[lanes % 16, lanes // 16] = a_lanes[ks]
            B[lanes // 16, lanes % 16] = bf[ks]
            acc += A @ B
        # D[4v + l/16][l%16] held by lane l, item v
        rows = np.array([[4 * vv + l // 16 for vv in range(4)] for l in lanes])
        cols = np.array([[l % 16] * 4 for l in lanes])
        vals = acc[rows, cols]
        return rows, cols, vals
    acc = np.zeros((4, 4, 4))  # block, i, j
    b4 = (lanes >> 2) & 3
    for ks in range(a_lanes.shape[0]):
        A = np.zeros((4, 4, 4))
        B = np.zeros((4, 4, 4))
        A[b4, lanes & 3, lanes >> 4] = a_lanes[ks]
        B[b4, lanes >> 4, lanes & 3] = bf[ks]
        acc += A @ B
    # lane l holds D_b[l>>4][l&3] -> cell (y + (l>>4), x + 4b + (l&3)) = (y + (l>>4), x + (l & 15))
    rows = (lanes >> 4)[:, None]
    cols = (lanes & 15)[:, None]
    vals = acc[b4, lanes >> 4, lanes & 3][:, None]
    return rows, cols, vals


def emulate(v, layout, dense, x):
    """Replays one launch of variant v (v['kf'] fused steps) on padded array x."""
    R, kf, by, bx, px = v["R"], v["kf"], v["by"], v["bx"], v["px"]
    tab = ktable(v["mf"], R, v["shape"])
    bf = bfrag(v["mf"], R, v["shape"], dense)
    offs = lane_offsets(v, tab)
    EY, EX = by + 2 * kf * R, bx + 2 * kf * R
    BUF = EY * px
    nbuf = 2 if kf > 1 else 1
    assert v["lds"] == nbuf * BUF * 8
    assert v["lds"] <= 160 * 1024
    ext, org, n = layout.ext, layout.org, layout.n
    pitch = ext[2]
    out = np.zeros(layout.cells)
    ty_n, tx_n = -(-n[1] // by), -(-n[2] // bx)
    ntiles = ty_n * tx_n
    per = -(-ntiles // 8)
    grid = per * 8
    seen = set()
    BH = 16 if v["mf"] == 16 else 4
    for wg in range(grid):
        tile = (wg % 8) * per + wg // 8
        if tile >= ntiles:
            continue
        assert tile not in seen
        seen.add(tile)
        Y0 = (tile // tx_n) * by - kf * R
        X0 = (tile % tx_n) * bx - kf * R
        ay0, ax0 = org[1] + Y0, org[2] + X0
        lds = np.full(nbuf * BUF, np.nan)
        for idx in range(BUF):
            r, c = divmod(idx, px)
            ay, ax = ay0 + r, ax0 + c
            val = 0.0
            if c < EX and 0 <= ay < ext[1] and 0 <= ax < pitch:
                gi = ay * pitch + ax
                assert 0 <= gi < layout.cells
                val = x[gi]
            lds[idx] = val
            if kf > 1:
                lds[BUF + idx] = val
        full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
        for s in range(kf):
            src = (s & 1) * BUF
            dst = ((s + 1) & 1) * BUF
            H, W = by + 2 * R * (kf - 1 - s), bx + 2 * R * (kf - 1 - s)
            base = R * (s + 1)
            nby, nbx = -(-H // BH), -(-W // 16)
            new_vals = []
            for blk in range(nby * nbx):
                ba, bb = divmod(blk, nbx)
                y = base + min(BH * ba, H - BH)
                xx = base + min(16 * bb, W - 16)
                idx = src + y * px + xx + offs
                assert idx.min() >= src and idx.max() < src + BUF, (v["name"], s, y, xx)
                a = lds[idx]
                assert np.isfinite(a).all(), (v["name"], "reads an unwritten LDS cell")
                rows, cols, vals = mfma_block(v, a, bf)
                for rr, cc, vv in zip(rows.reshape(-1), cols.reshape(-1), vals.reshape(-1)):
                    ty, tx = y + rr, xx + cc
                    inside = full or (0 <= Y0 + ty < n[1] and 0 <= X0 + tx < n[2])
                    if not inside:
                        continue
                    if s < kf - 1:
                        li = dst + ty * px + tx
                        assert dst <= li < dst + BUF
                        new_vals.append((li, vv))
                    else:
                        gi = (ay0 + ty) * pitch + ax0 + tx
                        assert 0 <= gi < layout.cells
                        assert org[1] <= ay0 + ty < org[1] + n[1] and org[2] <= ax0 + tx < org[2] + n[2]
                        out[gi] = vv
            for li, vv in new_vals:  # the kernel's barrier separates reads of src from these writes
                lds[li] = vv
    return out


def _dense(R, shape, seed):
    rng = np.random.default_rng(seed)
    D = 2 * R + 1
    w = rng.uniform(0.05, 1.0, (D, D))
    if shape == 1:
        mask = np.zeros((D, D), dtype=bool)
        mask[R, :] = True
        mask[:, R] = True
        w[~mask] = 0.0
    w /= w.sum()
    return w.reshape(-1)


VARIANTS = variants()


@pytest.mark.parametrize("v", VARIANTS, ids=[v["name"] for v in VARIANTS])
def test_variant_emulation(v):
    R, kf = v["R"], v["kf"]
    dense = _dense(R, v["shape"], v["id"])
    # partial tiles on both axes, and the reference's padded layout (extra columns, origin != halo)
    for layout in [Layout.with_halo([v["by"] + 5, v["bx"] + 19], R),
                   Layout(2, [v["by"] // 2 + 2 * R + 3, v["bx"] + 2 * R + 2 + 7], [R + 1, R + 2],
                          [v["by"] // 2, v["bx"] + 7])]:
        x = oracle.fill_data(layout, v["id"])
        got = emulate(v, layout, dense, x)
        offs, w = oracle.dense_to_taps(2, R, dense)
        want = oracle.run(layout, offs, w, x, kf)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], err)


def test_ktables_cover_every_tap():
    # every (tap, output column) pair appears exactly once in each table
    for mf in (16, 4):
        for R in (1, 2, 3):
            for shape in (0, 1):
                dy, c0, live = ktable(mf, R, shape)
                taps = {(a, b) for a in range(-R, R + 1) for b in range(-R, R + 1)
                        if shape == 0 or a == 0 or b == 0}
                width = 16 if mf == 16 else 4
                if mf == 16:
                    cols = [(dy[k], c0[k] + kk) for k in range(len(dy)) for kk in range(4)]
                else:
                    cols = [(dy[k], c0[k]) for k in range(len(dy)) if live[k]]
                    assert len(dy) % 4 == 0
                for (tdy, tdx) in taps:
                    for n in range(width):
                        hits = [1 for (a, c) in cols if a == tdy and c - n == tdx]
                        assert len(hits) == 1, (mf, R, shape, tdy, tdx, n)
