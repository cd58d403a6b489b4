"""CPU emulation of the 2-D FP64 MFMA kernels (k2d_f64.hip) -- no GPU needed.

For every compiled variant this replays the kernel's tile loop, LDS addressing,
block coverage, halo masking and MFMA lane algebra in numpy, with:
  * the k-step tables and B fragments taken from the library itself
    (include/convstencil_debug.h), so the tables that run on the GPU are checked;
  * the MFMA operand/result lane layouts measured on gfx950 (see the layout test
    in tests/test_gpu_mfma_layout.py):
      v_mfma_f64_16x16x4: A[l%16][l/16], B[l/16][l%16], D[4v + l/16][l%16]
      v_mfma_f64_4x4x4:   block (l>>2)&3, A[l&3][l>>4], B[l>>4][l&3], D[l>>4][l&3]
                          (2-D: block b's D[m][n] is cell (2b + (n>>1), 2m + (n&1)) of
                          an 8 x 8 output block, the dilated-table mapping);
  * an assertion on every LDS and global index (in bounds of the buffers the
    kernel allocates), so an out-of-bounds access is found here, not on a GPU;
  * the result compared with the CPU oracle (1e-13 relative).
"""
import ctypes
import math

import numpy as np
import pytest

from convstencil_amd import _lib
from convstencil_amd.shapes import Layout, dense_weights, shape_offsets
from oracle import oracle


def _dbg():
    lib = _lib.load()
    lib.cs_debug_variant_count.restype = ctypes.c_int32
    lib.cs_debug_variant_count.argtypes = [ctypes.c_int32]
    lib.cs_debug_variant_info.restype = ctypes.c_int32
    lib.cs_debug_variant_info.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_char_p, ctypes.c_int32]
    lib.cs_debug_bfrag_1d.restype = ctypes.c_int32
    lib.cs_debug_bfrag_1d.argtypes = [ctypes.c_int32, _lib.c_double_p, _lib.c_double_p, ctypes.c_int32]
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_ktable.restype = ctypes.c_int32
    lib.cs_debug_ktable.argtypes = [ctypes.c_int32] * 3 + [ip, ip, ip, ctypes.c_int32]
    lib.cs_debug_bfrag_2d.restype = ctypes.c_int32
    lib.cs_debug_bfrag_2d.argtypes = [ctypes.c_int32] * 3 + [_lib.c_double_p, _lib.c_double_p, ctypes.c_int32]
    return lib


def variants(ndim=2):
    lib = _dbg()
    out = []
    for i in range(lib.cs_debug_variant_count(ndim)):
        info = (ctypes.c_int64 * 12)()
        name = ctypes.create_string_buffer(128)
        assert lib.cs_debug_variant_info(ndim, i, info, name, 128) == 0
        keys = ["id", "mf", "R", "shape", "kf", "by", "bx", "nw", "lds", "px", "st", "cs"]
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


def dtable(R, shape):
    """Dilated 4x4x4 table: (np, ns, pp, ty[e], tx[e], live[e], map) with e = 4 * step + q."""
    lib = _dbg()
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_dtable.restype = ctypes.c_int32
    lib.cs_debug_dtable.argtypes = [ctypes.c_int32, ctypes.c_int32, ip, ip, ip, ip, ctypes.c_int32]
    cap = 256
    hdr = (ctypes.c_int32 * 4)()
    ty, tx, live = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
    n = lib.cs_debug_dtable(R, shape, hdr, ty, tx, live, cap)
    assert n > 0 and n == 4 * (hdr[0] + hdr[1])
    return hdr[0], hdr[1], hdr[2], np.array(ty[:n]).reshape(-1, 4), np.array(tx[:n]).reshape(-1, 4), \
        np.array(live[:n]).reshape(-1, 4), hdr[3]


def dil_row(dmap, b, i):
    """Row pair of the 2 x 2 sub-block of (block b, A/D row i) in an 8 x 8 block."""
    return i if dmap == 1 else b


def dil_col(dmap, b, i):
    return (b + i) & 3 if dmap == 1 else i


def lane_offsets(v, tab):
    """Per-lane LDS offsets off[read][lane] relative to the block's (y, x) cell."""
    px = v["px"]
    lanes = np.arange(64)
    if v["mf"] == 16:
        dy, c0, live = tab
        return np.array([((lanes & 15) + dy[ks]) * px + c0[ks] + (lanes >> 4) for ks in range(len(dy))])
    npair, ns, pp, ty, tx, live = tab[:6]
    dmap = tab[6] if len(tab) > 6 else 0
    q, b, i = lanes >> 4, (lanes >> 2) & 3, lanes & 3
    return np.array([(2 * dil_row(dmap, b, i) + ty[e][q]) * px + 2 * dil_col(dmap, b, i) + tx[e][q] + v["cs"]
                     for e in range(npair + ns)])


def mfma_block(v, a_lanes, bf, dmap=0):
    """a_lanes[ks][lane], bf[ks][lane] -> (rows, cols, vals) of the block's outputs per lane."""
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
    acc = np.zeros((4, 4, 4))  # block, m, n
    b4 = (lanes >> 2) & 3
    for ks in range(a_lanes.shape[0]):
        A = np.zeros((4, 4, 4))
        B = np.zeros((4, 4, 4))
        A[b4, lanes & 3, lanes >> 4] = a_lanes[ks]
        B[b4, lanes >> 4, lanes & 3] = bf[ks]
        acc += A @ B
    # lane l = 16 m + 4 b + n holds D_b[m][n]: cell (2 row + (n >> 1), 2 col + (n & 1)) of the
    # 2 x 2 sub-block (row, col) = (dil_row, dil_col)(b, m)
    n = lanes & 3
    rows = (2 * dil_row(dmap, b4, lanes >> 4) + (n >> 1))[:, None]
    cols = (2 * dil_col(dmap, b4, lanes >> 4) + (n & 1))[:, None]
    vals = acc[b4, lanes >> 4, n][:, None]
    return rows, cols, vals


def a_operands(v, tab, lds, idx):
    """The A operand of every k-step, as the kernel feeds it: pair reads give two k-steps."""
    if v["mf"] == 16:
        return lds[idx]
    npair, ns = tab[0], tab[1]
    out = []
    for e in range(npair):  # 16-byte read at idx[e] (aligned): .x then .y
        assert np.all(idx[e] % 2 == 0), (v["name"], "misaligned pair read")
        out.append(lds[idx[e]])
        out.append(lds[idx[e] + 1])
    for e in range(npair, npair + ns):
        out.append(lds[idx[e]])
    return np.array(out)


def _tile_geometry(v):
    R, kf, by, bx, px = v["R"], v["kf"], v["by"], v["bx"], v["px"]
    EY, EX = by + 2 * kf * R, bx + 2 * kf * R
    nbuf = 2 if (kf > 1 and v["st"] == 0) else 1
    return EY, EX, EY * px, nbuf


def _tile_load(v, layout, x, ay0, ax0):
    """The LDS image load_tile / load_tile_full give a tile: cells of the array, zero
    outside it and in pad columns."""
    EY, EX, BUF, nbuf = _tile_geometry(v)
    px, CS = v["px"], v["cs"]
    ext = layout.ext
    pitch = ext[2]
    lds = np.full(nbuf * BUF, np.nan)
    for idx in range(BUF):
        r, c = divmod(idx, px)
        c -= CS
        ay, ax = ay0 + r, ax0 + c
        val = 0.0
        if 0 <= c < EX and 0 <= ay < ext[1] and 0 <= ax < pitch:
            gi = ay * pitch + ax
            assert 0 <= gi < layout.cells
            val = x[gi]
        lds[idx] = val
        if nbuf == 2:
            lds[BUF + idx] = val
    return lds


def _tile_steps(v, tab, bf, offs, layout, lds, tile, out):
    """Replays the KF fused steps of one tile whose input is in lds (one buffer image,
    modified in place) and writes its output block to out."""
    R, kf, by, bx, px = v["R"], v["kf"], v["by"], v["bx"], v["px"]
    CS = v["cs"]
    EY, EX, BUF, nbuf = _tile_geometry(v)
    pp = tab[2] if v["mf"] == 4 else 0
    ext, org, n = layout.ext, layout.org, layout.n
    pitch = ext[2]
    tx_n = -(-n[2] // bx)
    Y0 = (tile // tx_n) * by - kf * R
    X0 = (tile % tx_n) * bx - kf * R
    ay0, ax0 = org[1] + Y0, org[2] + X0
    BH = BW = 16 if v["mf"] == 16 else 8
    full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
    for s in range(kf):
        # ST = 0 ping-pongs two buffers; ST = 1 reads and (after a barrier) rewrites one
        src = (s & 1) * BUF if nbuf == 2 else 0
        dst = ((s + 1) & 1) * BUF if nbuf == 2 else 0
        H, W = by + 2 * R * (kf - 1 - s), bx + 2 * R * (kf - 1 - s)
        base = R * (s + 1)
        if v["mf"] == 16:
            lox, lastx = base, base + W - 16
        else:  # block starts with x + CS + pp even, the first possibly one column early
            lox = base - ((base + CS + pp) & 1)
            lastx = (base + W - 8) + (((base + W - 8) + CS + pp) & 1)
        nby, nbx = -(-H // BH), -(-(lastx - lox) // BW) + 1
        new_vals = []
        for blk in range(nby * nbx):
            ba, bb = divmod(blk, nbx)
            y = base + min(BH * ba, H - BH)
            xx = min(lox + BW * bb, lastx)
            idx = src + y * px + xx + offs
            # a pair read's second cell; single reads (after the pairs) end at their cell
            top = max(idx[:tab[0]].max() + 1, idx.max()) if v["mf"] == 4 else idx.max()
            assert idx.min() >= src and top < src + BUF, (v["name"], s, y, xx)
            # every LDS row read stays inside the row (no wrap into the next row)
            a = a_operands(v, tab, lds, idx)
            assert np.isfinite(a).all(), (v["name"], "reads an unwritten LDS cell")
            rows, cols, vals = mfma_block(v, a, bf, tab[6] if v["mf"] == 4 else 0)
            ty, tx, vv = y + rows.reshape(-1), xx + cols.reshape(-1), vals.reshape(-1)
            inside = np.ones(ty.shape, dtype=bool) if full else \
                (Y0 + ty >= 0) & (Y0 + ty < n[1]) & (X0 + tx >= 0) & (X0 + tx < n[2])
            to_lds = s < kf - 1 or v["st"] == 2  # ST 2 stages the last step in LDS too
            if v["mf"] == 16 or not to_lds:
                # 16x16x4, and every step's grid writes: only the step's region columns
                inside &= (tx >= base) & (tx < base + W)
            # 4x4x4 LDS writes are not limited to the region: cells a block computes
            # outside it are never read by a later step's region
            if to_lds:
                li = dst + ty * px + tx + CS
                assert np.all(li >= dst) and np.all(li < dst + BUF) and np.all(ty < EY), (v["name"], s)
                new_vals.append((li[inside], vv[inside]))
            else:
                gi = (ay0 + ty) * pitch + ax0 + tx
                g = gi[inside]
                assert np.all(g >= 0) and np.all(g < layout.cells)
                assert np.all((ay0 + ty[inside] >= org[1]) & (ay0 + ty[inside] < org[1] + n[1]))
                assert np.all((ax0 + tx[inside] >= org[2]) & (ax0 + tx[inside] < org[2] + n[2]))
                out[g] = vv[inside]
        for li, vv in new_vals:  # the kernel's barrier separates reads of src from these writes
            lds[li] = vv
    if v["st"] == 2:  # coalesced copy-out of the output block from LDS
        off0 = kf * R
        for idx in range(by * bx):
            r, c = off0 + idx // bx, off0 + idx % bx
            if full or (Y0 + r < n[1] and X0 + c < n[2]):
                li = r * px + c + CS
                assert 0 <= li < BUF and np.isfinite(lds[li])
                gi = (ay0 + r) * pitch + ax0 + c
                assert org[1] <= ay0 + r < org[1] + n[1] and org[2] <= ax0 + c < org[2] + n[2]
                out[gi] = lds[li]


def _tables(v, dense):
    R = v["R"]
    tab = ktable(v["mf"], R, v["shape"]) if v["mf"] == 16 else dtable(R, v["shape"])
    bf = bfrag(v["mf"], R, v["shape"], dense)
    offs = lane_offsets(v, tab)
    if v["mf"] == 4:
        assert v["px"] % 2 == 0
        assert bf.shape[0] == 2 * tab[0] + tab[1]
    return tab, bf, offs


def emulate(v, layout, dense, x):
    """Replays one launch of variant v (v['kf'] fused steps) on padded array x."""
    R, kf, by, bx = v["R"], v["kf"], v["by"], v["bx"]
    tab, bf, offs = _tables(v, dense)
    EY, EX, BUF, nbuf = _tile_geometry(v)
    assert v["lds"] == (nbuf * BUF + 2) * 8  # two spare cells after the buffers
    assert v["lds"] <= 160 * 1024
    org, n = layout.org, layout.n
    out = np.zeros(layout.cells)
    ty_n, tx_n = -(-n[1] // by), -(-n[2] // bx)
    ntiles = ty_n * tx_n
    per = -(-ntiles // 8)
    grid = per * 8
    seen = set()
    for wg in range(grid):
        tile = (wg % 8) * per + wg // 8
        if tile >= ntiles:
            continue
        assert tile not in seen
        seen.add(tile)
        Y0 = (tile // tx_n) * by - kf * R
        X0 = (tile % tx_n) * bx - kf * R
        lds = _tile_load(v, layout, x, org[1] + Y0, org[2] + X0)
        _tile_steps(v, tab, bf, offs, layout, lds, tile, out)
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
    # plus 3 x 3 tiles, whose centre tile lies inside the interior (the unmasked path)
    for layout in [Layout.with_halo([v["by"] + 5, v["bx"] + 19], R),
                   Layout(2, [v["by"] // 2 + 2 * R + 3, v["bx"] + 2 * R + 2 + 7], [R + 1, R + 2],
                          [v["by"] // 2, v["bx"] + 7]),
                   Layout.with_halo([3 * v["by"] - 3, 3 * v["bx"] - 5], R)]:
        x = oracle.fill_data(layout, v["id"])
        got = emulate(v, layout, dense, x)
        offs, w = oracle.dense_to_taps(2, R, dense)
        want = oracle.run(layout, offs, w, x, kf)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], err)


def test_dtables_cover_every_tap():
    # every (tap, output cell) pair of an 8 x 8 block appears exactly once among the live
    # k-slots of the dilated table: slot t serves output v_n through tap t - v_n
    for R in (1, 2, 3):
        for shape in (0, 1):
            npair, ns, pp, ty, tx, live, dmap = dtable(R, shape)
            slots = []
            for e in range(4 * npair):
                if live[e // 4][e % 4]:
                    assert (tx[e // 4][e % 4] - pp) % 2 == 0  # aligned pair start
                    slots += [(ty[e // 4][e % 4], tx[e // 4][e % 4]), (ty[e // 4][e % 4], tx[e // 4][e % 4] + 1)]
            for e in range(4 * npair, 4 * (npair + ns)):
                if live[e // 4][e % 4]:
                    slots.append((ty[e // 4][e % 4], tx[e // 4][e % 4]))
            taps = {(a, b) for a in range(-R, R + 1) for b in range(-R, R + 1) if shape == 0 or a == 0 or b == 0}
            for (tdy, tdx) in taps:
                for n in range(4):
                    vy, vx = n >> 1, n & 1
                    assert slots.count((tdy + vy, tdx + vx)) == 1, (R, shape, tdy, tdx, n)
            # k-slot counts: the footprint size, rounded up to whole k-steps
            foot = {(a + vy, b + vx) for (a, b) in taps for vy in (0, 1) for vx in (0, 1)}
            assert 4 * (2 * npair + ns) < len(foot) + 8


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


# ---------------------------------------------------------------- 1-D family

def emulate_1d(v, layout, w, x):
    R, kf, L, P = v["R"], v["kf"], v["bx"], v["px"]
    lib = _dbg()
    cap = 64 * 16
    bfr = np.zeros(cap)
    wd = np.ascontiguousarray(w, dtype=np.float64)
    nks = lib.cs_debug_bfrag_1d(R, wd.ctypes.data_as(_lib.c_double_p), bfr.ctypes.data_as(_lib.c_double_p), cap)
    assert nks > 0
    bf = bfr[:nks * 64].reshape(nks, 64)
    lanes = np.arange(64)
    k, b4, i = lanes >> 4, (lanes >> 2) & 3, lanes & 3
    ne = 4 + 2 * R
    offs = []
    for ks in range(nks):
        f = 4 * ks + k
        c = np.where(f < ne, f - R, 0)
        offs.append(16 * i + 4 * b4 + c)
    offs = np.array(offs)
    E = L + 2 * kf * R
    assert P == E + 1
    nbuf = 2 if kf > 1 else 1
    assert v["lds"] == (nbuf * P + 2) * 8 and v["lds"] <= 160 * 1024
    ext, org, n = layout.ext[2], layout.org[2], layout.n[2]
    out = np.zeros(layout.cells)
    for tile in range(-(-n // L)):
        X0 = tile * L - kf * R
        ax0 = org + X0
        lds = np.full(nbuf * P, np.nan)
        for idx in range(P):
            ax = ax0 + idx
            val = x[ax] if (idx < E and 0 <= ax < ext) else 0.0
            lds[idx] = val
            if kf > 1:
                lds[P + idx] = val
        full = X0 >= 0 and X0 + E <= n
        for s in range(kf):
            src, dst = (s & 1) * P, ((s + 1) & 1) * P
            W = L + 2 * R * (kf - 1 - s)
            base = R * (s + 1)
            pend = []
            for blk in range(-(-W // 64)):
                xx = base + min(64 * blk, W - 64)
                idx = src + xx + offs
                assert idx.min() >= src and idx.max() < src + P
                a = lds[idx]
                assert np.isfinite(a).all()
                acc = np.zeros((4, 4, 4))
                for ks in range(nks):
                    A = np.zeros((4, 4, 4))
                    B = np.zeros((4, 4, 4))
                    A[b4, i, k] = a[ks]
                    B[b4, k, lanes & 3] = bf[ks]
                    acc += A @ B
                vals = acc[b4, lanes >> 4, lanes & 3]
                for ln in range(64):
                    tx = xx + ln
                    if not (full or 0 <= X0 + tx < n):
                        continue
                    if s < kf - 1:
                        pend.append((dst + tx, vals[ln]))
                    else:
                        gi = ax0 + tx
                        assert org <= gi < org + n
                        out[gi] = vals[ln]
            for li, vv in pend:
                lds[li] = vv
    return out


VARIANTS_1D = variants(1)


@pytest.mark.parametrize("v", VARIANTS_1D, ids=[v["name"] for v in VARIANTS_1D])
def test_variant_emulation_1d(v):
    R, kf = v["R"], v["kf"]
    rng = np.random.default_rng(v["id"])
    w = rng.uniform(0.05, 1.0, 2 * R + 1)
    w /= w.sum()
    for layout in [Layout.with_halo([v["bx"] + 77], R), Layout(1, [v["bx"] // 2 + 2 * R + 9], [R + 4],
                                                                  [v["bx"] // 2 + 3])]:
        x = oracle.fill_data(layout, v["id"])
        got = emulate_1d(v, layout, w, x)
        offs = [(0, 0, d) for d in range(-R, R + 1)]
        want = oracle.run(layout, offs, w, x, kf)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], err)


def emulate_1d_reg(v, layout, w, x):
    """k1d_reg: wave segments of 64 P cells (lane l holds cells l P .. l P + P - 1), KF steps
    in registers with the R cells beyond each lane's run from lanes l -+ 1 (zero past lanes 0
    and 63), then the central 64 P - 2 R KF cells stored.  Counts the stores per cell."""
    R, kf, P, NW = v["R"], v["kf"], v["px"], v["nw"]
    H = R * kf
    O = 64 * P - 2 * H
    assert v["mf"] == 6 and v["bx"] == NW * O and v["lds"] == 0 and O > 0 and P >= R
    lib = _dbg()
    cap = 64 * 16
    bfr = np.zeros(cap)
    wd = np.ascontiguousarray(w, dtype=np.float64)
    nks = lib.cs_debug_bfrag_1d(R, wd.ctypes.data_as(_lib.c_double_p), bfr.ctypes.data_as(_lib.c_double_p), cap)
    assert nks > 0
    # the kernel's weight reads: entry (k-row f, column 0) of the 4 x 4 x 4 B fragment
    wk = np.array([bfr[(f >> 2) * 64 + ((f & 3) << 4)] for f in range(2 * R + 1)])
    assert np.array_equal(wk, wd)
    ext, org, n = layout.ext[2], layout.org[2], layout.n[2]
    out = np.zeros(layout.cells)
    hits = np.zeros(layout.cells, dtype=np.int64)
    blocks = -(-n // v["bx"])
    d = np.arange(64 * P)
    for seg in range(blocks * NW):
        if seg * O >= n:
            continue
        X0 = seg * O - H
        c = X0 + d
        a = org + c
        ok = (a >= 0) & (a < ext)
        vals = np.where(ok, x[np.where(ok, a, 0)], 0.0)
        full = X0 >= 0 and X0 + 64 * P <= n
        for _ in range(kf):
            pad = np.concatenate([np.zeros(R), vals, np.zeros(R)])
            new = wk[0] * pad[0:64 * P]
            for t in range(1, 2 * R + 1):
                new = new + wk[t] * pad[t:t + 64 * P]
            if not full:
                new = np.where((c >= 0) & (c < n), new, vals)
            vals = new
        st = (d >= H) & (d < H + O) & (c < n)
        assert (c[st] >= 0).all() and (a[st] < org + n).all()
        out[a[st]] = vals[st]
        hits[a[st]] += 1
    assert (layout.interior(hits) == 1).all()
    return out


VARIANTS_1D_REG = variants(7)


@pytest.mark.parametrize("v", VARIANTS_1D_REG, ids=[v["name"] for v in VARIANTS_1D_REG])
def test_variant_emulation_1d_reg(v):
    R, kf = v["R"], v["kf"]
    rng = np.random.default_rng(v["id"])
    w = rng.uniform(0.05, 1.0, 2 * R + 1)
    w /= w.sum()
    O = v["bx"] // v["nw"]
    for layout in [Layout.with_halo([v["bx"] + 77], R), Layout(1, [O // 2 + 2 * R + 9], [R + 4], [O // 2 + 3]),
                   Layout.with_halo([3 * O + 5], R + 2)]:
        x = oracle.fill_data(layout, v["id"])
        got = emulate_1d_reg(v, layout, w, x)
        offs = [(0, 0, d) for d in range(-R, R + 1)]
        want = oracle.run(layout, offs, w, x, kf)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], err)


# ---------------------------------------------------------------- 3-D family

def ktable3d(R, shape):
    lib = _dbg()
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_ktable3d.restype = ctypes.c_int32
    lib.cs_debug_ktable3d.argtypes = [ctypes.c_int32, ctypes.c_int32, ip, ip, ip, ip, ctypes.c_int32]
    cap = 256
    dz = (ctypes.c_int32 * cap)()
    dy, c, live = (ctypes.c_int32 * (4 * cap))(), (ctypes.c_int32 * (4 * cap))(), (ctypes.c_int32 * (4 * cap))()
    n = lib.cs_debug_ktable3d(R, shape, dz, dy, c, live, cap)
    assert n > 0
    return (np.array(dz[:n]), np.array(dy[:4 * n]).reshape(n, 4), np.array(c[:4 * n]).reshape(n, 4),
            np.array(live[:4 * n]).reshape(n, 4))


def bfrag3d(R, shape, w):
    lib = _dbg()
    lib.cs_debug_bfrag_3d.restype = ctypes.c_int32
    lib.cs_debug_bfrag_3d.argtypes = [ctypes.c_int32, ctypes.c_int32, _lib.c_double_p, _lib.c_double_p,
                                      ctypes.c_int32]
    cap = 64 * 256
    out = np.zeros(cap)
    wd = np.ascontiguousarray(w, dtype=np.float64)
    nks = lib.cs_debug_bfrag_3d(R, shape, wd.ctypes.data_as(_lib.c_double_p), out.ctypes.data_as(_lib.c_double_p),
                                cap)
    assert nks > 0
    return out[:nks * 64].reshape(nks, 64)


def emulate_3d(v, layout, dense, x):
    R, BY, BX, ZC, nw = v["R"], v["by"], v["bx"], v["st"], v["nw"]
    dz_t, dy_t, c_t, live_t = ktable3d(R, v["shape"])
    bf = bfrag3d(R, v["shape"], dense)
    EY, EX = BY + 2 * R, BX + 2 * R
    PX = v["px"]
    assert PX >= EX
    PLANE = EY * PX
    NP = 2 * R + 1
    assert v["lds"] == NP * PLANE * 8 and v["lds"] <= 160 * 1024
    ext, org, n = layout.ext, layout.org, layout.n
    x3 = x.reshape(ext)
    out = np.zeros(layout.cells)
    o3 = out.reshape(ext)
    tiles = (-(-n[0] // ZC), -(-n[1] // BY), -(-n[2] // BX))
    lanes = np.arange(64)
    k, b4, i4 = lanes >> 4, (lanes >> 2) & 3, lanes & 3
    for zc in range(tiles[0]):
        for ty in range(tiles[1]):
            for tx in range(tiles[2]):
                Y0, X0 = ty * BY - R, tx * BX - R
                ay0, ax0 = org[1] + Y0, org[2] + X0
                Z0 = zc * ZC
                Z1 = min(Z0 + ZC, n[0])
                ring = np.full((NP, PLANE), np.nan)

                def load(p):
                    pl = np.zeros(PLANE)
                    az = org[0] + p
                    assert 0 <= az < ext[0]
                    for idx in range(PLANE):
                        r, c = divmod(idx, PX)
                        ay, ax = ay0 + r, ax0 + c
                        if c < EX and ay < ext[1] and ax < ext[2]:
                            assert ay >= 0 and ax >= 0
                            pl[idx] = x3[az, ay, ax]
                    return pl
                slot = lambda p: (p - (Z0 - R)) % NP
                for p in range(Z0 - R, Z0 + R + 1):
                    ring[slot(p)] = load(p)
                for z in range(Z0, Z1):
                    for blk in range((BY // 4) * (BX // 16)):
                        y, xx = R + 4 * (blk // (BX // 16)), R + 16 * (blk % (BX // 16))
                        acc = np.zeros((4, 4, 4))
                        for ks in range(len(dz_t)):
                            pl = ring[slot(z + dz_t[ks])]
                            idx = (y + i4 + dy_t[ks][k]) * PX + xx + 4 * b4 + c_t[ks][k]
                            assert idx.min() >= 0 and idx.max() < PLANE
                            a = pl[idx]
                            assert np.isfinite(a).all()
                            A = np.zeros((4, 4, 4))
                            B = np.zeros((4, 4, 4))
                            A[b4, i4, k] = a
                            B[b4, k, lanes & 3] = bf[ks]
                            acc += A @ B
                        vals = acc[b4, lanes >> 4, lanes & 3]
                        for ln in range(64):
                            tyy, txx = y + (ln >> 4), xx + (ln & 15)
                            if Y0 + tyy < n[1] and X0 + txx < n[2]:
                                gi = (org[0] + z, ay0 + tyy, ax0 + txx)
                                assert all(org[d] <= gi[d] < org[d] + n[d] for d in range(3))
                                o3[gi] = vals[ln]
                    if z + 1 < Z1:
                        ring[slot(z + R + 1)] = load(z + R + 1)
    return out


def dtable3(R, shape, dz):
    lib = _dbg()
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_dtable3.restype = ctypes.c_int32
    lib.cs_debug_dtable3.argtypes = [ctypes.c_int32] * 3 + [ip, ip, ip, ip, ctypes.c_int32]
    cap = 256
    hdr = (ctypes.c_int32 * 3)()
    ty, tx, live = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
    n = lib.cs_debug_dtable3(R, shape, dz, hdr, ty, tx, live, cap)
    assert n > 0 and n == 4 * (hdr[0] + hdr[1])
    return hdr[0], hdr[1], hdr[2], np.array(ty[:n]).reshape(-1, 4), np.array(tx[:n]).reshape(-1, 4), \
        np.array(live[:n]).reshape(-1, 4)


def bfrag3d2(R, shape, w):
    lib = _dbg()
    lib.cs_debug_bfrag_3d2.restype = ctypes.c_int32
    lib.cs_debug_bfrag_3d2.argtypes = [ctypes.c_int32, ctypes.c_int32, _lib.c_double_p, _lib.c_double_p,
                                       ctypes.c_int32]
    cap = 64 * 256
    out = np.zeros(cap)
    wd = np.ascontiguousarray(w, dtype=np.float64)
    nks = lib.cs_debug_bfrag_3d2(R, shape, wd.ctypes.data_as(_lib.c_double_p), out.ctypes.data_as(_lib.c_double_p),
                                 cap)
    assert nks > 0
    return out[:nks * 64].reshape(nks, 64)


def emulate_3d2(v, layout, dense, x):
    """The 8 x 8 block 3-D kernel (mfma shape 8): ring of 2R+1 planes, dilated table per dz."""
    R, BY, BX, ZC = v["R"], v["by"], v["bx"], v["st"]
    tabs = {dz: dtable3(R, v["shape"], dz) for dz in range(-R, R + 1)}
    pps = {t[2] for t in tabs.values()}
    assert len(pps) == 1
    pp = pps.pop()
    bf = bfrag3d2(R, v["shape"], dense)
    assert bf.shape[0] == sum(2 * t[0] + t[1] for t in tabs.values())
    EY, EX, PX, CS = BY + 2 * R, BX + 2 * R, v["px"], v["cs"]
    assert PX % 2 == 0
    PLANE = EY * PX
    NP = 2 * R + 1
    assert v["lds"] == (NP * PLANE + 2) * 8 and v["lds"] <= 160 * 1024
    lanes = np.arange(64)
    q, b4, m4 = lanes >> 4, (lanes >> 2) & 3, lanes & 3
    orow, ocol = 2 * b4 + ((lanes & 3) >> 1), 2 * (lanes >> 4) + (lanes & 1)
    LOX = R - ((R + CS + pp) & 1)
    LASTX = (R + BX - 8) + (((R + BX - 8) + CS + pp) & 1)
    NBX = (LASTX - LOX + 7) // 8 + 1
    ext, org, n = layout.ext, layout.org, layout.n
    x3 = x.reshape(ext)
    out = np.zeros(layout.cells)
    o3 = out.reshape(ext)
    tiles = (-(-n[0] // ZC), -(-n[1] // BY), -(-n[2] // BX))
    for zc in range(tiles[0]):
        for ty_ in range(tiles[1]):
            for tx_ in range(tiles[2]):
                Y0, X0 = ty_ * BY - R, tx_ * BX - R
                ay0, ax0 = org[1] + Y0, org[2] + X0
                Z0, Z1 = zc * ZC, min(zc * ZC + ZC, n[0])
                ring = np.full((NP, PLANE), np.nan)

                def load(p):
                    pl = np.zeros(PLANE)
                    az = org[0] + p
                    assert 0 <= az < ext[0]
                    for idx in range(PLANE):
                        r, c = divmod(idx, PX)
                        c -= CS
                        ay, ax = ay0 + r, ax0 + c
                        if 0 <= c < EX and 0 <= ay < ext[1] and 0 <= ax < ext[2]:
                            pl[idx] = x3[az, ay, ax]
                    return pl
                slot = lambda p: (p - (Z0 - R)) % NP
                for p in range(Z0 - R, Z0 + R + 1):
                    ring[slot(p)] = load(p)
                for z in range(Z0, Z1):
                    for blk in range((BY // 8) * NBX):
                        ba, bb = divmod(blk, NBX)
                        y, xx = R + 8 * ba, min(LOX + 8 * bb, LASTX)
                        acc = np.zeros((4, 4, 4))
                        ks = 0
                        for dz in range(-R, R + 1):
                            npair, ns, _, tty, ttx, tlive = tabs[dz]
                            pl = ring[slot(z + dz)]
                            for e in range(npair + ns):
                                idx = (y + 2 * b4 + tty[e][q]) * PX + xx + 2 * m4 + ttx[e][q] + CS
                                assert idx.min() >= 0 and idx.max() + (1 if e < npair else 0) < PLANE
                                if e < npair:
                                    assert np.all(idx % 2 == 0), "misaligned pair read"
                                    cols = [idx, idx + 1]
                                else:
                                    cols = [idx]
                                for ci in cols:
                                    a = pl[ci]
                                    assert np.isfinite(a).all()
                                    A = np.zeros((4, 4, 4))
                                    B = np.zeros((4, 4, 4))
                                    A[b4, m4, q] = a
                                    B[b4, q, lanes & 3] = bf[ks]
                                    acc += A @ B
                                    ks += 1
                        vals = acc[b4, lanes >> 4, lanes & 3]
                        for ln in range(64):
                            tyy, txx = y + orow[ln], xx + ocol[ln]
                            if R <= txx < R + BX and Y0 + tyy < n[1] and X0 + txx < n[2]:
                                gi = (org[0] + z, ay0 + tyy, ax0 + txx)
                                assert all(org[d] <= gi[d] < org[d] + n[d] for d in range(3))
                                o3[gi] = vals[ln]
                    if z + 1 < Z1:
                        ring[slot(z + R + 1)] = load(z + R + 1)
    return out


def emulate_3f(v, layout, dense, x):
    """The fused 3-D kernel: KF levels streamed through z, a ring of 2R+1 planes per level."""
    R, BY, BX, ZC, KF = v["R"], v["by"], v["bx"], v["st"], v["kf"]
    tabs = {dz: dtable3(R, v["shape"], dz) for dz in range(-R, R + 1)}
    pp = tabs[0][2]
    bf = bfrag3d2(R, v["shape"], dense)
    EY, EX, PX, CS = BY + 2 * KF * R, BX + 2 * KF * R, v["px"], v["cs"]
    assert PX % 2 == 0
    PLANE = EY * PX
    NP = 2 * R + 1
    assert v["lds"] == (KF * NP * PLANE + 2) * 8 and v["lds"] <= 160 * 1024
    lanes = np.arange(64)
    q4, b4, m4 = lanes >> 4, (lanes >> 2) & 3, lanes & 3
    orow, ocol = 2 * b4 + ((lanes & 3) >> 1), 2 * (lanes >> 4) + (lanes & 1)

    def geo(l):
        LO, H = R * l, BY + 2 * R * (KF - l)
        HI = LO + BX + 2 * R * (KF - l)
        LOX = LO - ((LO + CS + pp) & 1)
        LASTX = (HI - 8) + (((HI - 8) + CS + pp) & 1)
        return LO, H, HI, LOX, LASTX, -(-H // 8), (LASTX - LOX + 7) // 8 + 1
    ext, org, n = layout.ext, layout.org, layout.n
    x3 = x.reshape(ext)
    out = np.zeros(layout.cells)
    o3 = out.reshape(ext)
    tiles = (-(-n[0] // ZC), -(-n[1] // BY), -(-n[2] // BX))
    for zc in range(tiles[0]):
        for ty_ in range(tiles[1]):
            for tx_ in range(tiles[2]):
                Y0, X0 = ty_ * BY - KF * R, tx_ * BX - KF * R
                ay0, ax0 = org[1] + Y0, org[2] + X0
                Z0, Z1 = zc * ZC, min(zc * ZC + ZC, n[0])
                base = Z0 - KF * R
                full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
                rings = np.full((KF, NP, PLANE), np.nan)
                rings[1:] = 0.0
                slot = lambda p: (p - base) % NP

                def load(p):
                    pl = np.zeros(PLANE)
                    az = org[0] + p
                    if not (0 <= az < ext[0]):
                        return pl
                    for idx in range(PLANE):
                        r, c = divmod(idx, PX)
                        c -= CS
                        ay, ax = ay0 + r, ax0 + c
                        if 0 <= c < EX and 0 <= ay < ext[1] and 0 <= ax < ext[2]:
                            pl[idx] = x3[az, ay, ax]
                    return pl

                def inside(ty, tx):
                    return full or (0 <= Y0 + ty < n[1] and 0 <= X0 + tx < n[2])
                for p in range(base, Z1 - 1 + KF * R + 1):
                    rings[0, slot(p)] = load(p)
                    for l in range(1, KF + 1):
                        qz = p - l * R
                        if not (Z0 - (KF - l) * R <= qz <= Z1 - 1 + (KF - l) * R):
                            continue
                        if l < KF and (qz < 0 or qz >= n[0]):
                            rings[l, slot(qz)] = 0.0
                            continue
                        LO, H, HI, LOX, LASTX, NBY, NBX = geo(l)
                        writes = []
                        for blk in range(NBY * NBX):
                            ba, bb = divmod(blk, NBX)
                            y, xx = LO + min(8 * ba, H - 8), min(LOX + 8 * bb, LASTX)
                            acc = np.zeros((4, 4, 4))
                            ks = 0
                            for dz in range(-R, R + 1):
                                npair, ns, _, tty, ttx, tlive = tabs[dz]
                                pl = rings[l - 1, slot(qz + dz)]
                                for e in range(npair + ns):
                                    idx = (y + 2 * b4 + tty[e][q4]) * PX + xx + 2 * m4 + ttx[e][q4] + CS
                                    assert idx.min() >= 0 and idx.max() + (1 if e < npair else 0) < PLANE
                                    assert e >= npair or np.all(idx % 2 == 0), "misaligned pair read"
                                    for ci in ([idx, idx + 1] if e < npair else [idx]):
                                        a = pl[ci]
                                        assert np.isfinite(a).all(), (v["name"], "reads an unset LDS cell")
                                        A = np.zeros((4, 4, 4))
                                        B = np.zeros((4, 4, 4))
                                        A[b4, m4, q4] = a
                                        B[b4, q4, lanes & 3] = bf[ks]
                                        acc += A @ B
                                        ks += 1
                            vals = acc[b4, lanes >> 4, lanes & 3]
                            for ln in range(64):
                                tyy, txx = y + orow[ln], xx + ocol[ln]
                                if LO <= txx < HI and inside(tyy, txx):
                                    writes.append((tyy, txx, vals[ln]))
                        for tyy, txx, val in writes:
                            if l < KF:
                                rings[l, slot(qz), tyy * PX + txx + CS] = val
                            else:
                                gi = (org[0] + qz, ay0 + tyy, ax0 + txx)
                                assert all(org[d] <= gi[d] < org[d] + n[d] for d in range(3))
                                o3[gi] = val
    return out


VARIANTS_3D = variants(3)


@pytest.mark.parametrize("v", VARIANTS_3D, ids=[v["name"] for v in VARIANTS_3D])
def test_variant_emulation_3d(v):
    R, sh = v["R"], v["shape"]
    D = 2 * R + 1
    rng = np.random.default_rng(v["id"])
    w = rng.uniform(0.05, 1.0, (D, D, D))
    zz, yy, xx = np.meshgrid(*[np.arange(-R, R + 1)] * 3, indexing="ij")
    if sh == 1:
        w[((zz != 0).astype(int) + (yy != 0) + (xx != 0)) > 1] = 0.0
    elif sh == 2:
        w[np.abs(zz) + np.abs(yy) + np.abs(xx) > R] = 0.0
    w /= w.sum()
    dense = w.reshape(-1)
    layout = Layout(3, [v["st"] + 2 * R + 4, v["by"] + 2 * R + 5, v["bx"] // 2 + 2 * R + 9], [R + 1, R, R + 2],
                    [v["st"] + 3, v["by"] + 5, v["bx"] // 2 + 7])
    x = oracle.fill_data(layout, v["id"])
    if v["kf"] > 1:  # fused kernels run with a zero halo only
        m = layout.interior_mask()
        x = np.where(m, x, 0.0)
        got = emulate_3f(v, layout, dense, x)
    else:
        got = emulate_3d2(v, layout, dense, x) if v["mf"] == 8 else emulate_3d(v, layout, dense, x)
    offs, ww = oracle.dense_to_taps(3, R, dense)
    want = oracle.run(layout, offs, ww, x, v["kf"])
    gi, wi = layout.interior(got), layout.interior(want)
    err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
    assert err < 1e-13, (v["name"], err)


# ---------------------------------------------------------------- 2-D split precision

SPLIT = {1: ("bf16", 8, 32, 3, 3, 3, 1.0), 2: ("f32", 1, 4, 2, 2, 2, 1.0), 3: ("f16", 8, 32, 2, 2, 2, 2048.0),
         4: ("f16", 8, 32, 2, 3, 3, 2048.0)}


def keep_product(prec, s, q):
    """Slice products (state slice s, weight slice q) the dilated split form keeps below
    level 0: every level below NL, except w1 u1 for FP16 with three weight slices."""
    NSU, NSW, NL = SPLIT[prec][3], SPLIT[prec][4], SPLIT[prec][5]
    if not 1 <= s + q < NL:
        return False
    return s == 0 or q == 0 if (NSU, NSW) == (2, 3) else True


def _to_f32(raw, kind):
    if kind == "bf16":
        return (raw.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    if kind == "f16":
        return raw.view(np.float16).astype(np.float32)
    return raw.view(np.float32)


def _round(x, kind):
    x32 = np.asarray(x, dtype=np.float64).astype(np.float32)
    if kind == "bf16":
        b = x32.view(np.uint32).astype(np.uint64)
        return (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    if kind == "f16":
        return x32.astype(np.float16).astype(np.float32)
    return x32


def split_np(v, kind, ns, scale):
    rem = np.asarray(v, dtype=np.float64)
    out, sc = [], 1.0
    for _ in range(ns):
        t = _round(rem * sc, kind)
        out.append(t)
        rem = rem - t.astype(np.float64) / sc
        sc *= scale
    return out


def emulate_split(v, layout, dense, x):
    prec, R, KF, BY, BX = v["mf"], v["R"], v["kf"], v["by"], v["bx"]
    kind, CK, KI, NSU, NSW, NL, SCALE = SPLIT[prec]
    esz = 4 if kind == "f32" else 2
    lib = _dbg()
    lib.cs_debug_bfrag_split.restype = ctypes.c_int64
    lib.cs_debug_bfrag_split.argtypes = [ctypes.c_int32, ctypes.c_int32, _lib.c_double_p, ctypes.c_void_p,
                                         ctypes.c_int64]
    raw = np.zeros(1 << 20, dtype=np.uint8)
    wd = np.ascontiguousarray(dense, dtype=np.float64)
    nbytes = lib.cs_debug_bfrag_split(prec, R, wd.ctypes.data_as(_lib.c_double_p), raw.ctypes.data, raw.size)
    assert nbytes > 0
    el = _to_f32(raw[:nbytes].copy().view(np.uint32 if esz == 4 else np.uint16), kind)
    RPD = 24 if CK == 8 else ((16 + 2 * R + 3) // 4) * 4
    NENT = (2 * R + 1) * RPD
    NINS = -(-NENT // KI)
    bf = el.reshape(NSW, NINS, 64, CK)
    EY, EX = BY + 2 * KF * R, BX + 2 * KF * R
    TOP = (((R - KF * R) % 8) + 8) % 8 + 8 if CK == 8 else 8
    PR = -(-(TOP + EY + RPD - 2 * R - 1) // 8) * 8
    assert v["px"] == PR
    PLANE = EX * PR
    assert v["lds"] == (2 * NSU * PLANE + 8) * esz and v["lds"] <= 160 * 1024
    lanes = np.arange(64)
    ai, q = lanes & 15, lanes >> 4

    def ent(e):
        return (e // RPD - R, e % RPD) if e < NENT else (0, 0)
    off = np.zeros((NINS, 64), dtype=np.int64)
    for m in range(NINS):
        for ln in range(64):
            dx, rr = ent(m * KI + q[ln] * CK)
            off[m, ln] = (ai[ln] + dx) * PR + rr
    ext, org, n = layout.ext, layout.org, layout.n
    pitch = ext[2]
    out = np.zeros(layout.cells)
    x3 = x.reshape(ext)
    ty_n, tx_n = -(-n[1] // BY), -(-n[2] // BX)
    for tile in range(ty_n * tx_n):
        Y0, X0 = (tile // tx_n) * BY - KF * R, (tile % tx_n) * BX - KF * R
        ay0, ax0 = org[1] + Y0, org[2] + X0
        full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
        vals = np.zeros(PLANE)
        for idx in range(PLANE):
            c, sr = divmod(idx, PR)
            r = sr - TOP
            if 0 <= r < EY:
                ay, ax = ay0 + r, ax0 + c
                if 0 <= ay < ext[1] and 0 <= ax < pitch:
                    vals[idx] = x3[0, ay, ax]
        sl = split_np(vals, kind, NSU, SCALE)
        bufs = [np.stack(sl), np.stack(sl)]
        for st in range(KF):
            src, dst = bufs[st & 1], bufs[(st + 1) & 1]
            last = st == KF - 1
            lo, hi = R * (st + 1), EY - R * (st + 1)
            W = EX - 2 * R * (st + 1)
            y0 = lo - (((lo - KF * R) % 8) + 8) % 8
            nby, nbx = -(-(hi - y0) // 16), -(-W // 16)
            writes = []
            for blk in range(nby * nbx):
                ba, bb = divmod(blk, nbx)
                y, xx = y0 + 16 * ba, lo + min(16 * bb, W - 16)
                ubase = xx * PR + (y - R + TOP)
                acc = [np.zeros((16, 16), dtype=np.float32) for _ in range(NL)]
                for m in range(NINS):
                    base = ubase + off[m]
                    if CK == 8:
                        assert np.all(base % 8 == 0), "16-byte LDS read misaligned"
                    assert base.min() >= 0 and base.max() + CK <= PLANE
                    for s in range(NSU):
                        a = src[s][base[:, None] + np.arange(CK)[None, :]]  # [lane][e]
                        A = np.zeros((16, KI), dtype=np.float32)
                        A[ai[:, None], q[:, None] * CK + np.arange(CK)[None, :]] = a
                        for t in range(NSW):
                            if s + t >= NL:
                                continue
                            B = np.zeros((KI, 16), dtype=np.float32)
                            B[q[:, None] * CK + np.arange(CK)[None, :], ai[:, None]] = bf[t, m]
                            acc[s + t] = (acc[s + t] + (A.astype(np.float64) @ B.astype(np.float64)).astype(
                                np.float32)).astype(np.float32)
                for i in range(16):        # output column x + i
                    for nn in range(16):   # output row y + nn
                        tyy, txx = y + nn, xx + i
                        val, sc = 0.0, 1.0
                        for L in range(NL):
                            val += float(acc[L][i, nn]) / sc
                            sc *= SCALE
                        if not (lo <= tyy < hi):
                            continue
                        if not (full or (0 <= Y0 + tyy < n[1] and 0 <= X0 + txx < n[2])):
                            continue
                        if last:
                            gi = (ay0 + tyy) * pitch + ax0 + txx
                            assert org[1] <= ay0 + tyy < org[1] + n[1] and org[2] <= ax0 + txx < org[2] + n[2]
                            out[gi] = val
                        else:
                            writes.append((txx * PR + tyy + TOP, val))
            for li, val in writes:
                assert 0 <= li < PLANE
                for k, s_ in enumerate(split_np(np.array([val]), kind, NSU, SCALE)):
                    dst[k][li] = s_[0]
    return out


VARIANTS_SPLIT = variants(4)


VARIANTS_SPLIT1 = [v for v in VARIANTS_SPLIT if v["st"] == 0 and v["kf"] <= 2]


@pytest.mark.parametrize("v", VARIANTS_SPLIT1, ids=[v["name"] for v in VARIANTS_SPLIT1])
def test_variant_emulation_split(v):
    R = v["R"]
    dense = _dense(R, 0, v["id"])
    layout = Layout(2, [v["by"] // 2 + 2 * R + 3, v["bx"] + 2 * R + 9], [R + 1, R + 2], [v["by"] // 2, v["bx"] + 7])
    x = oracle.fill_data(layout, v["id"])
    got = emulate_split(v, layout, dense, x)
    offs, w = oracle.dense_to_taps(2, R, dense)
    want = oracle.run(layout, offs, w, x, v["kf"])
    gi, wi = layout.interior(got), layout.interior(want)
    err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
    assert err < 2e-6, (v["name"], err)


# ------------------------------------------- 2-D split precision, dilated form (staging 2)

def ctable(R, shape):
    lib = _dbg()
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_ctable.restype = ctypes.c_int32
    lib.cs_debug_ctable.argtypes = [ctypes.c_int32, ctypes.c_int32, ip, ip, ip, ip, ctypes.c_int32]
    cap = 128
    c0 = ctypes.c_int32()
    ty, tx, live = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
    n = lib.cs_debug_ctable(R, shape, ctypes.byref(c0), ty, tx, live, cap)
    assert n > 0 and n % 4 == 0
    return c0.value, np.array(ty[:n]), np.array(tx[:n]), np.array(live[:n])


def _f32_to(x, kind):
    """FP32 -> slice type, round to nearest even (as the kernel's conversions), back as FP32."""
    return _round(np.asarray(x, dtype=np.float32), kind)


def split_f32_np(v32, kind, ns, scale):
    """The dilated split kernels' FP32 split (split4 / split_f32): three slices, each the
    remainder rounded to nearest (exact for BF16x3), or FP16x2's rounded high slice and
    scaled remainder."""
    v32 = np.asarray(v32, dtype=np.float32)
    if ns == 3:
        h = _f32_to(v32, kind)
        r1 = (v32 - h).astype(np.float32)
        m = _f32_to(r1, kind)
        lo = _f32_to((r1 - m).astype(np.float32), kind)
        return [h, m, lo]
    h = _f32_to(v32, kind)
    return [h, _f32_to(((v32 - h) * np.float32(scale)).astype(np.float32), kind)]


def emulate_split2(v, layout, dense, x):
    prec, R, SH, KF, BY, BX = v["mf"], v["R"], v["shape"], v["kf"], v["by"], v["bx"]
    kind, CK, KI, NSU, NSW, NL, SCALE = SPLIT[prec]
    assert kind in ("bf16", "f16")
    c0, cty, ctx, clive = ctable(R, SH)
    NINS = len(cty) // 4
    lib = _dbg()
    lib.cs_debug_bfrag_split2.restype = ctypes.c_int64
    lib.cs_debug_bfrag_split2.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _lib.c_double_p,
                                          ctypes.c_void_p, ctypes.c_int64]
    raw = np.zeros(1 << 20, dtype=np.uint8)
    wd = np.ascontiguousarray(dense, dtype=np.float64)
    nbytes = lib.cs_debug_bfrag_split2(prec, R, SH, wd.ctypes.data_as(_lib.c_double_p), raw.ctypes.data, raw.size)
    assert nbytes == NSW * NINS * 64 * 8 * 2
    wf = _to_f32(raw[:nbytes].copy().view(np.uint16), kind).reshape(NSW, NINS, 64, 8)
    EY, EX, PX, CS = BY + 2 * KF * R, BX + 2 * KF * R, v["px"], v["cs"]
    assert PX % 8 == 0
    PLANE = EY * PX
    assert v["lds"] == (NSU * PLANE + 8) * 2 and v["lds"] <= 160 * 1024
    lanes = np.arange(64)
    grp, nn = lanes >> 4, lanes & 15
    # A[m][k] from lane l: m = l & 15, k = 8 (l >> 4) + j;  B[k][n]: n = l & 15, k = 8 (l >> 4) + j
    A = np.zeros((NSW, NINS, 16, 32), dtype=np.float64)
    for q in range(NSW):
        for i in range(NINS):
            A[q, i][nn[:, None], 8 * grp[:, None] + np.arange(8)[None, :]] = wf[q, i]
    lib.cs_debug_ctable_map.restype = ctypes.c_int32
    lib.cs_debug_ctable_map.argtypes = [ctypes.c_int32, ctypes.c_int32]
    amap = lib.cs_debug_ctable_map(R, SH)
    arow = nn & 7
    aoct = ((nn >> 3) ^ ((nn >> 2) & 1)) if amap == 1 else (nn >> 3)
    off = np.array([(2 * arow + cty[4 * i + grp]) * PX + 8 * aoct + ctx[4 * i + grp] + CS
                    for i in range(NINS)])
    # D lane l holds m = 4 (l >> 4) + v, n = l & 15: cell (2 arow + (m >> 3), 8 aoct + (m & 7))
    orow = 2 * arow + (grp >> 1)
    ocol = 8 * aoct + 4 * (grp & 1)
    INV = np.float32(1.0 / SCALE)
    ext, org, n = layout.ext, layout.org, layout.n
    pitch = ext[2]
    out = np.zeros(layout.cells)
    x3 = x.reshape(ext)
    ty_n, tx_n = -(-n[1] // BY), -(-n[2] // BX)
    for tile in range(ty_n * tx_n):
        Y0, X0 = (tile // tx_n) * BY - KF * R, (tile % tx_n) * BX - KF * R
        ay0, ax0 = org[1] + Y0, org[2] + X0
        full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
        vals = np.zeros(PLANE)
        for idx in range(PLANE):
            r, c = divmod(idx, PX)
            c -= CS
            ay, ax = ay0 + r, ax0 + c
            if 0 <= c < EX and 0 <= ay < ext[1] and 0 <= ax < pitch:
                vals[idx] = x3[0, ay, ax]
        # the FP64 tile is rounded to FP32 once and split like the step results
        lds = [s_.astype(np.float32) for s_ in split_f32_np(vals.astype(np.float32), kind, NSU, SCALE)]
        staged = ",os1" in v["name"]
        ob = np.full(BY * BX, np.nan)

        def inside(ty, tx):
            return full or (0 <= Y0 + ty < n[1] and 0 <= X0 + tx < n[2])
        for st in range(KF):
            last = st == KF - 1
            H, W = BY + 2 * R * (KF - 1 - st), BX + 2 * R * (KF - 1 - st)
            LO = R * (st + 1)
            HI = LO + W
            LOX = LO - ((LO + CS + c0) % 8)
            nby, nbx = -(-H // 16), -(-(HI - LOX) // 16)
            writes = []
            for blk in range(nby * nbx):
                ba, bb = divmod(blk, nbx)
                y, xx = LO + min(16 * ba, H - 16), LOX + 16 * bb
                d0 = np.zeros((16, 16), dtype=np.float32)
                d1 = np.zeros((16, 16), dtype=np.float32)
                for i in range(NINS):
                    base = y * PX + xx + off[i]
                    assert np.all(base % 8 == 0), (v["name"], "16-byte LDS read misaligned")
                    assert base.min() >= 0 and base.max() + 8 <= PLANE
                    assert np.all(base % PX + 8 <= PX), (v["name"], "read wraps into the next row")
                    B = []
                    for s in range(NSU):
                        g = lds[s][base[:, None] + np.arange(8)[None, :]]  # [lane][j]
                        Bs = np.zeros((32, 16), dtype=np.float64)
                        Bs[8 * grp[:, None] + np.arange(8)[None, :], nn[:, None]] = g
                        B.append(Bs)
                    d0 = (d0 + (A[0, i] @ B[0]).astype(np.float32)).astype(np.float32)
                    for s in range(NSU):
                        for q in range(NSW):
                            if keep_product(prec, s, q):
                                d1 = (d1 + (A[q, i] @ B[s]).astype(np.float32)).astype(np.float32)
                for ln in range(64):
                    ty = y + orow[ln]
                    for vv in range(4):
                        m = 4 * grp[ln] + vv
                        tx = xx + ocol[ln] + vv
                        a0, a1 = d0[m, nn[ln]], d1[m, nn[ln]]
                        if last and staged:
                            if LO <= tx < HI:  # compact BY x BX FP64 block over the slice planes
                                oi = (ty - LO) * BX + tx - LO
                                assert 0 <= oi < BY * BX and (oi + 1) * 8 <= v["lds"]
                                ob[oi] = float(a0) + float(a1) * float(INV)
                        elif last:
                            if LO <= tx < HI and inside(ty, tx):
                                gi = (ay0 + ty) * pitch + ax0 + tx
                                assert org[1] <= ay0 + ty < org[1] + n[1] and org[2] <= ax0 + tx < org[2] + n[2]
                                out[gi] = float(a0) + float(a1) * float(INV)
                        else:
                            all4 = full or (inside(ty, xx + ocol[ln]) and inside(ty, xx + ocol[ln] + 3))
                            if all4 or inside(ty, tx):
                                li = ty * PX + tx + CS
                                assert 0 <= li < PLANE and 0 <= tx + CS < PX
                                writes.append((li, np.float32(a0 + a1 * INV)))
            for li, val in writes:
                for k, s_ in enumerate(split_f32_np(np.array([val], dtype=np.float32), kind, NSU, SCALE)):
                    lds[k][li] = s_[0]
        if staged:  # coalesced copy-out of the staged block
            off0 = KF * R
            for idx in range(BY * BX):
                r, c = off0 + idx // BX, off0 + idx % BX
                if full or (Y0 + r < n[1] and X0 + c < n[2]):
                    assert np.isfinite(ob[idx])
                    gi = (ay0 + r) * pitch + ax0 + c
                    assert org[1] <= ay0 + r < org[1] + n[1] and org[2] <= ax0 + c < org[2] + n[2]
                    out[gi] = ob[idx]
    return out


VARIANTS_SPLIT2 = [v for v in VARIANTS_SPLIT if v["st"] == 2 and (v["kf"] <= 2 or ",os1" in v["name"])]


@pytest.mark.parametrize("v", VARIANTS_SPLIT2, ids=[v["name"] for v in VARIANTS_SPLIT2])
def test_variant_emulation_split2(v):
    R = v["R"]
    dense = _dense(R, v["shape"], v["id"])
    # the reference's padded layout (partial tiles), and 3 x 3 tiles whose centre tile
    # lies inside the interior (the unmasked path)
    for layout in [Layout(2, [v["by"] // 2 + 2 * R + 3, v["bx"] + 2 * R + 9], [R + 1, R + 2],
                          [v["by"] // 2, v["bx"] + 7]),
                   Layout.with_halo([3 * v["by"] - 3, 3 * v["bx"] - 5], R)]:
        x = oracle.fill_data(layout, v["id"])
        got = emulate_split2(v, layout, dense, x)
        offs, w = oracle.dense_to_taps(2, R, dense)
        want = oracle.run(layout, offs, w, x, v["kf"])
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 2e-6, (v["name"], err)
        # halo cells of the output are never written
        assert np.all(got[~layout.interior_mask()] == 0.0)


def test_ctables_cover_every_tap():
    # every (tap, output offset v) pair lands in exactly one live chunk slot: slot t of a
    # chunk serves output offset v_m = (m >> 3, m & 7) through tap t - v_m
    for R in (1, 2, 3):
        for shape in (0, 1):
            c0, ty, tx, live = ctable(R, shape)
            assert c0 in (0, 4) and np.all((tx - c0) % 8 == 0)
            slots = [(ty[e], tx[e] + j) for e in range(len(ty)) if live[e] for j in range(8)]
            assert len(slots) == len(set(slots))
            taps = {(a, b) for a in range(-R, R + 1) for b in range(-R, R + 1) if shape == 0 or a == 0 or b == 0}
            for (tdy, tdx) in taps:
                for m in range(16):
                    assert (tdy + (m >> 3), tdx + (m & 7)) in slots, (R, shape, tdy, tdx, m)


# ------------------------------------------------------------- 3-D split precision

def ctable3s(R, shape, dz):
    lib = _dbg()
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.cs_debug_ctable3.restype = ctypes.c_int32
    lib.cs_debug_ctable3.argtypes = [ctypes.c_int32] * 3 + [ip, ip, ip, ip, ctypes.c_int32]
    cap = 128
    c0 = ctypes.c_int32()
    ty, tx, live = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
    n = lib.cs_debug_ctable3(R, shape, dz, ctypes.byref(c0), ty, tx, live, cap)
    assert n > 0 and n % 4 == 0
    return c0.value, np.array(ty[:n]), np.array(tx[:n]), np.array(live[:n])


def emulate_split3(v, layout, dense, x):
    """The split-precision 3-D kernel (k3d_split.hip): KF levels streamed through z, a ring
    of 2R+1 planes of slices per level, 16 x 16 blocks with one chunk table per plane."""
    import re
    prec, R, SH, KF, BY, BX = v["mf"], v["R"], v["shape"], v["kf"], v["by"], v["bx"]
    ZC = int(re.search(r"zc(\d+)", v["name"]).group(1))
    kind, CK, KI, NSU, NSW, NL, SCALE = SPLIT[prec]
    tabs = {dz: ctable3s(R, SH, dz) for dz in range(-R, R + 1)}
    c0 = tabs[0][0]
    assert all(t[0] == c0 for t in tabs.values())
    nins = {dz: len(tabs[dz][1]) // 4 for dz in tabs}
    ib = {dz: sum(nins[d] for d in range(-R, dz)) for dz in tabs}
    NINS = sum(nins.values())
    lib = _dbg()
    lib.cs_debug_bfrag_split3.restype = ctypes.c_int64
    lib.cs_debug_bfrag_split3.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _lib.c_double_p,
                                          ctypes.c_void_p, ctypes.c_int64]
    raw = np.zeros(1 << 20, dtype=np.uint8)
    wd = np.ascontiguousarray(dense, dtype=np.float64)
    nbytes = lib.cs_debug_bfrag_split3(prec, R, SH, wd.ctypes.data_as(_lib.c_double_p), raw.ctypes.data, raw.size)
    assert nbytes == NSW * NINS * 64 * 8 * 2
    wf = _to_f32(raw[:nbytes].copy().view(np.uint16), kind).reshape(NSW, NINS, 64, 8)
    EY, EX, PX, CS = BY + 2 * KF * R, BX + 2 * KF * R, v["px"], v["cs"]
    assert PX % 8 == 0
    PLANE, NP = EY * PX, 2 * R + 1
    assert v["lds"] == (KF * NP * NSU * PLANE + 8) * 2 and v["lds"] <= 160 * 1024
    lanes = np.arange(64)
    grp, nn = lanes >> 4, lanes & 15
    A = np.zeros((NSW, NINS, 16, 32), dtype=np.float64)
    for q in range(NSW):
        for i in range(NINS):
            A[q, i][nn[:, None], 8 * grp[:, None] + np.arange(8)[None, :]] = wf[q, i]
    off = np.zeros((NINS, 64), dtype=np.int64)
    for dz in tabs:
        _, cty, ctx, _ = tabs[dz]
        for li in range(nins[dz]):
            off[ib[dz] + li] = (2 * (nn & 7) + cty[4 * li + grp]) * PX + 8 * (nn >> 3) + ctx[4 * li + grp] + CS
    orow = 2 * (nn & 7) + (grp >> 1)
    ocol = 8 * (nn >> 3) + 4 * (grp & 1)
    INV = np.float32(1.0 / SCALE)

    def geo(l):
        LO, H = R * l, BY + 2 * R * (KF - l)
        HI = LO + BX + 2 * R * (KF - l)
        LOX = LO - ((LO + CS + c0) % 8)
        return LO, H, HI, LOX, -(-H // 16), -(-(HI - LOX) // 16)
    ext, org, n = layout.ext, layout.org, layout.n
    x3 = x.reshape(ext)
    out = np.zeros(layout.cells)
    o3 = out.reshape(ext)
    tiles = (-(-n[0] // ZC), -(-n[1] // BY), -(-n[2] // BX))
    rr, cc = np.divmod(np.arange(PLANE), PX)
    cc = cc - CS
    for zc in range(tiles[0]):
        for ty_ in range(tiles[1]):
            for tx_ in range(tiles[2]):
                Y0, X0 = ty_ * BY - KF * R, tx_ * BX - KF * R
                ay0, ax0 = org[1] + Y0, org[2] + X0
                Z0, Z1 = zc * ZC, min(zc * ZC + ZC, n[0])
                base = Z0 - KF * R
                full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
                rings = np.full((KF, NP, NSU, PLANE), np.nan, dtype=np.float32)
                rings[1:] = 0.0
                slot = lambda p: (p - base) % NP

                def load(p):
                    pl = np.zeros(PLANE)
                    az = org[0] + p
                    if 0 <= az < ext[0]:
                        ay, ax = ay0 + rr, ax0 + cc
                        ok = (cc >= 0) & (cc < EX) & (ay >= 0) & (ay < ext[1]) & (ax >= 0) & (ax < ext[2])
                        pl[ok] = x3[az, ay[ok], ax[ok]]
                    return np.stack(split_f32_np(pl.astype(np.float32), kind, NSU, SCALE))

                def inside(ty, tx):
                    return full or (0 <= Y0 + ty < n[1] and 0 <= X0 + tx < n[2])
                for p in range(base, Z1 - 1 + KF * R + 1):
                    rings[0, slot(p)] = load(p)
                    for l in range(1, KF + 1):
                        qz = p - l * R
                        if not (Z0 - (KF - l) * R <= qz <= Z1 - 1 + (KF - l) * R):
                            continue
                        if l < KF and (qz < 0 or qz >= n[0]):
                            rings[l, slot(qz)] = 0.0
                            continue
                        LO, H, HI, LOX, NBY, NBX = geo(l)
                        writes = []
                        for blk in range(NBY * NBX):
                            ba, bb = divmod(blk, NBX)
                            y, xx = LO + min(16 * ba, H - 16), LOX + 16 * bb
                            d0 = np.zeros((16, 16), dtype=np.float32)
                            d1 = np.zeros((16, 16), dtype=np.float32)
                            for dz in range(-R, R + 1):
                                src = rings[l - 1, slot(qz + dz)]
                                for li in range(nins[dz]):
                                    i = ib[dz] + li
                                    b0 = y * PX + xx + off[i]
                                    assert np.all(b0 % 8 == 0), (v["name"], "16-byte LDS read misaligned")
                                    assert b0.min() >= 0 and b0.max() + 8 <= PLANE
                                    assert np.all(b0 % PX + 8 <= PX), (v["name"], "read wraps into the next row")
                                    B = []
                                    for s in range(NSU):
                                        g = src[s][b0[:, None] + np.arange(8)[None, :]]
                                        assert np.isfinite(g).all(), (v["name"], "reads an unset LDS cell")
                                        Bs = np.zeros((32, 16), dtype=np.float64)
                                        Bs[8 * grp[:, None] + np.arange(8)[None, :], nn[:, None]] = g
                                        B.append(Bs)
                                    d0 = (d0 + (A[0, i] @ B[0]).astype(np.float32)).astype(np.float32)
                                    for s in range(NSU):
                                        for q in range(NSW):
                                            if keep_product(prec, s, q):
                                                d1 = (d1 + (A[q, i] @ B[s]).astype(np.float32)).astype(np.float32)
                            for ln in range(64):
                                tyy = y + orow[ln]
                                for vv in range(4):
                                    m = 4 * grp[ln] + vv
                                    txx = xx + ocol[ln] + vv
                                    a0, a1 = d0[m, nn[ln]], d1[m, nn[ln]]
                                    if l == KF:
                                        if LO <= txx < HI and inside(tyy, txx):
                                            gi = (org[0] + qz, ay0 + tyy, ax0 + txx)
                                            assert all(org[d] <= gi[d] < org[d] + n[d] for d in range(3))
                                            o3[gi] = float(a0) + float(a1) * float(INV)
                                    elif inside(tyy, txx):  # full tiles: every cell, overshoot included
                                        li_ = tyy * PX + txx + CS
                                        assert 0 <= li_ < PLANE and 0 <= txx + CS < PX
                                        writes.append((li_, np.float32(a0 + a1 * INV)))
                        if l < KF:
                            for li_, val in writes:
                                for k, s_ in enumerate(split_f32_np(np.array([val], dtype=np.float32), kind, NSU,
                                                                    SCALE)):
                                    rings[l, slot(qz), k, li_] = s_[0]
    return out


VARIANTS_SPLIT3 = variants(5)


def _dense3(R, sh, seed):
    D = 2 * R + 1
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.05, 1.0, (D, D, D))
    zz, yy, xx = np.meshgrid(*[np.arange(-R, R + 1)] * 3, indexing="ij")
    if sh == 1:
        w[((zz != 0).astype(int) + (yy != 0) + (xx != 0)) > 1] = 0.0
    w /= w.sum()
    return w.reshape(-1)


def _check_split3(v, layout):
    R = v["R"]
    dense = _dense3(R, v["shape"], v["id"])
    x = oracle.fill_data(layout, v["id"])
    x = np.where(layout.interior_mask(), x, 0.0)  # fused kernels run with a zero halo
    got = emulate_split3(v, layout, dense, x)
    offs, ww = oracle.dense_to_taps(3, R, dense)
    want = oracle.run(layout, offs, ww, x, v["kf"])
    gi, wi = layout.interior(got), layout.interior(want)
    err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
    assert err < 2e-6, (v["name"], err)
    assert np.all(got[~layout.interior_mask()] == 0.0)


@pytest.mark.parametrize("v", VARIANTS_SPLIT3, ids=[v["name"] for v in VARIANTS_SPLIT3])
def test_variant_emulation_split3(v):
    # 3 x 3 tiles per plane: the centre tile lies inside the interior (unmasked path), the
    # others are cut by the domain edge (partial tiles); 7 planes, one z-chunk
    _check_split3(v, Layout.with_halo([7, 3 * v["by"] - 3, 3 * v["bx"] - 5], v["R"]))


def test_split3_zchunks():
    # two z-chunks (the second partial) for an unfused and a fused variant
    import re
    for kf in (1, 2):
        v = [u for u in VARIANTS_SPLIT3 if u["kf"] == kf][0]
        zc = int(re.search(r"zc(\d+)", v["name"]).group(1))
        _check_split3(v, Layout.with_halo([zc + 3, v["by"] + 5, v["bx"] // 2 + 7], v["R"]))


def test_ctable3s_cover_every_tap():
    # per plane: every (tap, output offset v) pair lands in exactly one live chunk slot
    for shape in (0, 1):
        R = 1
        for dz in range(-R, R + 1):
            c0, ty, tx, live = ctable3s(R, shape, dz)
            assert c0 in (0, 4) and np.all((tx - c0) % 8 == 0)
            slots = [(ty[e], tx[e] + j) for e in range(len(ty)) if live[e] for j in range(8)]
            assert len(slots) == len(set(slots))
            taps = {(a, b) for a in range(-R, R + 1) for b in range(-R, R + 1)
                    if shape == 0 or (dz == 0 and (a == 0 or b == 0)) or (a == 0 and b == 0)}
            for (tdy, tdx) in taps:
                for m in range(16):
                    vy, vx = m >> 3, m & 7
                    assert (tdy + vy, tdx + vx) in slots, (shape, dz, tdy, tdx, m)


def _bf16_rn(x32):
    """FP32 -> BF16 (round to nearest even), back as FP32."""
    u = np.asarray(x32, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def _plain_slices(dense, prec):
    """The library's plain weight split (split_value): each slice is the remainder rounded
    to FP32 and then to the slice type; FP16x2's second slice is scaled by 2^11."""
    rem = dense.copy()
    out = []
    if prec == 3:
        h = rem.astype(np.float32).astype(np.float16).astype(np.float64)
        l = ((rem - h) * 2048).astype(np.float32).astype(np.float16).astype(np.float64)
        return [h, l], 2048.0
    for _ in range(3):
        sl = _bf16_rn(rem.astype(np.float32)).astype(np.float64)
        out.append(sl)
        rem = rem - sl
    return out, 1.0


@pytest.mark.parametrize("prec", [3, 1])
def test_split_weights_compensated(prec):
    # the dilated form's weight slices (FP16x2: h + l / 2048; BF16x3: s0 + s1 + s2): the
    # last slice of every tap within 4 ulps of its plain split, the others equal to it, and
    # the errors d = w - (slices) cancelling in the sum and the first moments, which every
    # step repeats (a plain split leaves sums that compound over the steps)
    lib = _dbg()
    lib.cs_debug_bfrag_split2.restype = ctypes.c_int64
    lib.cs_debug_bfrag_split2.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _lib.c_double_p,
                                          ctypes.c_void_p, ctypes.c_int64]
    kind, ns = ("f16", 2) if prec == 3 else ("bf16", 3)
    for R, shape in ((3, 0), (3, 1), (1, 0), (1, 1)):
        D = 2 * R + 1
        dense = _dense(R, shape, 77 + R + shape).reshape(D, D)
        c0, cty, ctx, clive = ctable(R, shape)
        nins = len(cty) // 4
        raw = np.zeros(1 << 20, dtype=np.uint8)
        nbytes = lib.cs_debug_bfrag_split2(prec, R, shape, np.ascontiguousarray(dense.reshape(-1)).ctypes.data_as(
            _lib.c_double_p), raw.ctypes.data, raw.size)
        assert nbytes == ns * nins * 64 * 8 * 2
        wf = _to_f32(raw[:nbytes].copy().view(np.uint16), kind).astype(np.float64).reshape(ns, nins, 64, 8)
        sl = np.full((ns, D, D), np.nan)
        for i in range(nins):
            for lane in range(64):
                m, c = lane & 15, 4 * i + (lane >> 4)
                for j in range(8):
                    dy, dx = cty[c] - (m >> 3), ctx[c] + j - (m & 7)
                    if clive[c] and -R <= dy <= R and -R <= dx <= R:
                        for q in range(ns):
                            v = wf[q, i, lane, j]
                            assert np.isnan(sl[q, dy + R, dx + R]) or sl[q, dy + R, dx + R] == v
                            sl[q, dy + R, dx + R] = v
        assert not np.isnan(sl).any()
        plain, scale = _plain_slices(dense, prec)
        for q in range(ns - 1):
            assert np.array_equal(sl[q], plain[q]), (R, shape, q)
        lp = plain[-1]
        if prec == 3:
            ulp = np.spacing(np.abs(lp).astype(np.float16)).astype(np.float64)
        else:  # BF16: 2^(e - 7) for |v| in [2^e, 2^(e+1)); the smallest normal's for zero
            e = np.floor(np.log2(np.maximum(np.abs(lp), 2.0 ** -126)))
            ulp = 2.0 ** (e - 7)
        assert np.all(np.abs(sl[-1] - lp) <= 4 * ulp * (1 + 1e-12)), (R, shape)
        assert np.all((dense != 0) | (sl == 0).all(axis=0))
        yy, xx = np.meshgrid(np.arange(-R, R + 1), np.arange(-R, R + 1), indexing="ij")
        last_scale = scale ** (ns - 1)
        d = dense - (sum(sl[q] / scale ** q for q in range(ns)))
        dp = dense - (sum(plain[q] / scale ** q for q in range(ns)))
        assert last_scale > 0
        got = [abs(d.sum()), abs((d * yy).sum()), abs((d * xx).sum())]
        plain_m = [abs(dp.sum()), abs((dp * yy).sum()), abs((dp * xx).sum())]
        # the search stops at the granularity of the last slices' ulps: about 1e-10 for the
        # 49-tap box2d3r, coarser for 9 taps of about 1/9 (FP16x2 then models to about 1e-7
        # relative L2 after 100 box2d1r steps, BF16x3 to about 4e-8)
        assert got[0] <= max(0.1 * plain_m[0], 2e-10 if prec == 3 else 5e-10), (R, shape, got, plain_m)
        assert sum(got) <= max(0.25 * sum(plain_m), 4e-9 if prec == 3 else 1e-9), (R, shape, got, plain_m)


# ------------------------------------- 2-D FP64, pipelined persistent kernel (k2d_stream)

VARIANTS_STREAM = variants(6)


def emulate_stream(v, layout, dense, x, G):
    """Replays one launch of a pipelined persistent variant with G workgroups: each walks
    positions blockIdx.x, + G, ... of the XCD-ordered tile list, with two LDS tile buffers
    that keep their contents from tile to tile (the kernel zeroes them once).  The producer
    waves' direct copies are replayed dword by dword with their bounds: an interior tile
    copies its rows over whatever the buffer held (pad columns keep old values, which must
    stay finite), an edge tile is zeroed first and copies only the cells inside the array."""
    import re
    NP = int(re.search(r",p(\d+),", v["name"]).group(1))
    R, kf, by, bx, px, CS = v["R"], v["kf"], v["by"], v["bx"], v["px"], v["cs"]
    tab, bf, offs = _tables(v, dense)
    EY, EX, BUF, nbuf = _tile_geometry(v)
    assert nbuf == 1
    SLOT = BUF + 8
    assert v["lds"] == 2 * SLOT * 8 and v["lds"] <= 160 * 1024
    assert (v["nw"] + NP) * 64 <= 1024
    if v["st"] == 2:  # the staged copy-out covers the block in whole passes
        NT = v["nw"] * 64
        assert NT % bx == 0 and by % (NT // bx) == 0
    ext, org, n = layout.ext, layout.org, layout.n
    pitch = ext[2]
    out = np.zeros(layout.cells)
    tx_n = -(-n[2] // bx)
    ntiles = -(-n[1] // by) * tx_n
    per = -(-ntiles // 8)
    total = per * 8
    DW = 2 * EX
    CH = -(-DW // 64)
    TAIL = DW - 64 * (CH - 1)
    x32 = np.ascontiguousarray(x).view(np.uint32)
    seen = set()
    for wg in range(min(G, total)):
        bufs = [np.zeros(SLOT), np.zeros(SLOT)]
        cur = 0
        k = wg
        while k < total:
            tile = (k % 8) * per + k // 8
            if tile >= ntiles:
                k += G
                continue
            assert tile not in seen
            seen.add(tile)
            Y0 = (tile // tx_n) * by - kf * R
            X0 = (tile % tx_n) * bx - kf * R
            ay0, ax0 = org[1] + Y0, org[2] + X0
            full = Y0 >= 0 and X0 >= 0 and Y0 + EY <= n[1] and X0 + EX <= n[2]
            b = bufs[cur]
            if not full:
                b[:BUF] = 0.0
            b32 = b.view(np.uint32)
            for r in range(EY):
                ay = ay0 + r
                if not full and not (0 <= ay < ext[1]):
                    continue
                for c in range(CH):
                    lanes = np.arange(64 if c < CH - 1 else TAIL)
                    d = 64 * c + lanes
                    if not full:
                        axs = ax0 + (d >> 1)
                        lanes, d = lanes[(axs >= 0) & (axs < pitch)], d[(axs >= 0) & (axs < pitch)]
                    lb = (r * px + CS) * 8 + 4 * d  # LDS byte offsets
                    gb = (ay * pitch + ax0) * 8 + 4 * d  # global byte offsets
                    assert np.all(lb >= 0) and np.all(lb + 4 <= BUF * 8), (v["name"], r, c)
                    assert np.all(gb >= 0) and np.all(gb + 4 <= layout.cells * 8), (v["name"], r, c)
                    b32[lb // 4] = x32[gb // 4]
            _tile_steps(v, tab, bf, offs, layout, b, tile, out)
            cur ^= 1
            k += G
    assert len(seen) == ntiles
    return out


@pytest.mark.parametrize("v", VARIANTS_STREAM, ids=[v["name"] for v in VARIANTS_STREAM])
def test_variant_emulation_stream(v):
    R, kf = v["R"], v["kf"]
    dense = _dense(R, v["shape"], v["id"])
    for layout, G in [(Layout.with_halo([v["by"] + 5, v["bx"] + 19], R), 8),
                      (Layout(2, [v["by"] // 2 + 2 * R + 3, v["bx"] + 2 * R + 2 + 7], [R + 1, R + 2],
                              [v["by"] // 2, v["bx"] + 7]), 16),
                      (Layout.with_halo([3 * v["by"] - 3, 3 * v["bx"] - 5], R), 8),
                      (Layout.with_halo([4 * v["by"], 3 * v["bx"] + 1], R), 16)]:
        x = oracle.fill_data(layout, v["id"])
        got = emulate_stream(v, layout, dense, x, G)
        offs, w = oracle.dense_to_taps(2, R, dense)
        want = oracle.run(layout, offs, w, x, kf)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], err)


# ---------------------------------------------------------------- 2-D register-streaming family

def reg2d_split(ny, nx, ox, radius, kf, ring, slots):
    lib = _dbg()
    i64p = ctypes.POINTER(ctypes.c_int64)
    lib.cs_debug_reg2d_split.restype = ctypes.c_int32
    lib.cs_debug_reg2d_split.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, i64p]
    lib.cs_debug_reg2d_chunk.restype = ctypes.c_int32
    lib.cs_debug_reg2d_chunk.argtypes = [i64p, ctypes.c_int64, ctypes.c_int64, i64p]
    split = (ctypes.c_int64 * 5)()
    assert lib.cs_debug_reg2d_split(ny, nx, ox, radius, kf, ring, slots, split) == 0
    strips, he, hm, we, wi = list(split)
    ne = 2 if strips >= 2 else 1
    chunks = []
    for wid in range(ne * we + (strips - ne) * wi):
        rows = (ctypes.c_int64 * 3)()
        if lib.cs_debug_reg2d_chunk(split, ny, wid, rows) == 1:
            chunks.append(tuple(rows))
    return list(split), chunks


def reg_period(R, ring):
    ns = 2 * R + 1
    return ns // math.gcd(ns, ring) * ring


def emulate_2d_reg(v, layout, dense, x, slots):
    """k2d_reg: each wave sweeps rows [y0, y1) of a strip of 64 PX columns (lane l: columns
    l PX .. l PX + PX - 1, so that a cell's column neighbours are the nearby cells of the
    flat lane-major row, zero past the wave's ends), one input row per iteration in unrolled
    blocks of the period, rows loaded RING - 1 iterations ahead into a ring of RING rows,
    level s + 1 holding partial sums for rows yq - R .. yq + R in slots (y - ys) mod (2R + 1),
    the incoming row's taps of row offset dy = R .. -R added to row yq - dy (dy = R completes
    it, dy = -R starts it), cells outside the interior frozen on the checked path, level KF's
    completed row stored on the output columns.  Checks loads and stores against the array
    and counts the stores per cell."""
    R, PX, KF, RING = v["R"], v["px"], v["kf"], v["st"]
    NS, H = 2 * R + 1, R * KF
    PER = reg_period(R, RING)
    least = H + PER + RING - 2
    box = v["shape"] == 0
    OX = 64 * PX - 2 * H
    assert v["mf"] == 7 and v["bx"] == OX and v["lds"] == 0 and PER % NS == 0 and PER % RING == 0
    _, ey, pitch = layout.ext
    _, oy, ox = layout.org
    _, ny, nx = layout.n
    X = x.reshape(ey, pitch)
    D = 2 * R + 1
    w = np.asarray(dense, dtype=np.float64).reshape(D, D)
    out = np.zeros((ey, pitch))
    hits = np.zeros((ey, pitch), dtype=np.int64)
    _, chunks = reg2d_split(ny, nx, OX, R, KF, RING, slots)
    d = np.arange(64 * PX)
    plain = 0
    for strip, y0, y1 in chunks:
        xs = strip * OX - H
        cols = xs + d
        fast = xs >= 0 and xs + 64 * PX <= nx and y0 >= least and y1 + least <= ny
        plain += fast
        colin = (cols >= 0) & (cols < nx)
        colarr = (ox + cols >= 0) & (ox + cols < pitch)
        colout = (d >= H) & (d < H + OX) & (cols < nx)

        def load(r):
            ay = oy + r
            if fast:
                assert 0 <= ay < ey and colarr.all()
                return X[ay, ox + cols].copy()
            ok = colarr & (0 <= ay < ey)
            return np.where(ok, X[ay if 0 <= ay < ey else 0, np.where(ok, ox + cols, 0)], 0.0)

        ys, iters = y0 - H, (y1 - y0) + 2 * H
        acc = np.zeros((KF, NS, 64 * PX))
        ring = [load(ys + k) for k in range(RING - 1)] + [None]
        for it in range(0, iters, PER):
            for K in range(PER):
                r = ys + it + K
                ring[(K + RING - 1) % RING] = load(r + RING - 1)
                q = ring[K % RING].copy()
                for s in range(KF):
                    yq = r - R * s
                    qe = np.concatenate([np.zeros(R), q, np.zeros(R)])
                    for dy in range(R, -R - 1, -1):
                        sl = (K - R * s - dy) % NS
                        T = acc[s, sl]
                        t = None
                        for dx in range(-R, R + 1):
                            if not box and dy != 0 and dx != 0:
                                continue
                            term = w[dy + R, dx + R] * qe[R + dx:R + dx + 64 * PX]
                            t = term if (t is None and dy == -R) else (T if t is None else t) + term
                        if not fast:
                            yt = yq - dy
                            fr = (yt < 0 or yt >= ny) | ~colin
                            if dy > 0:
                                t = np.where(fr, T, t)
                            elif dy == 0:
                                t = np.where(fr, q, t)
                        acc[s, sl] = t
                    q = acc[s, (K - R * s - R) % NS].copy()
                yo = r - H
                if y0 <= yo < y1:
                    c = ox + cols[colout]
                    assert 0 <= oy + yo < ey and (c >= 0).all() and (c < pitch).all()
                    out[oy + yo, c] = q[colout]
                    hits[oy + yo, c] += 1
    inner = hits[oy:oy + ny, ox:ox + nx]
    assert (inner == 1).all(), (int(inner.min()), int(inner.max()))
    assert hits.sum() == inner.sum()
    emulate_2d_reg.plain_chunks = plain
    return out.reshape(-1)


VARIANTS_REG2D = variants(8)


@pytest.mark.parametrize("v", VARIANTS_REG2D, ids=[v["name"] for v in VARIANTS_REG2D])
def test_variant_emulation_2d_reg(v):
    # grids of one, two and three strips, with halos of 1 and of odd widths (masks at the
    # array's edges), few slots (tall chunks: the plain path inside) and many
    R, kf, OX = v["R"], v["kf"], v["bx"]
    name = {(1, 0): "box2d1r", (1, 1): "star2d1r", (3, 0): "box2d3r", (3, 1): "star2d3r"}[(R, v["shape"])]
    offs = shape_offsets(name)
    rng = np.random.default_rng(v["id"])
    w = rng.uniform(0.05, 1.0, len(offs))
    w /= w.sum()
    dense = dense_weights(2, R, offs, w)
    H = R * kf
    cases = [(Layout.with_halo([6 * H + 40, 2 * OX + 37], R), 6),
             (Layout(2, [4 * H + 31 + 2 * R, OX // 2 + 12 + 2 * R], [R + 1, R + 2], [4 * H + 26, OX // 2 + 5]), 3),
             (Layout.with_halo([9 * H + 25, OX + 11], R), 64),
             (Layout.with_halo([16 * H + 120, 3 * OX], R), 15)]
    for i, (layout, slots) in enumerate(cases):
        x = oracle.fill_data(layout, v["id"] + slots)
        got = emulate_2d_reg(v, layout, dense, x, slots)
        if i == 3:
            assert emulate_2d_reg.plain_chunks > 0  # the plain path ran
        want = oracle.run(layout, offs, w, x, kf)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], slots, err)


@pytest.mark.parametrize("ny,nx,ox,R,kf,ring,slots", [(10240, 10240, 104, 1, 12, 6, 2048),
                                                      (8192, 8192, 104, 3, 4, 7, 2048),
                                                      (8192, 8192, 240, 1, 8, 3, 2048), (100, 3000, 240, 1, 8, 3, 2048),
                                                      (5000, 200, 96, 1, 16, 6, 2048), (31, 31, 254, 1, 1, 3, 4096),
                                                      (1, 7, 240, 1, 8, 6, 16), (50, 70, 122, 3, 1, 7, 512)])
def test_reg2d_split_covers_once(ny, nx, ox, R, kf, ring, slots):
    # every row of every strip in exactly one chunk, never more waves than slots when the
    # strips fit
    split, chunks = reg2d_split(ny, nx, ox, R, kf, ring, slots)
    strips, he, hm, we, wi = split
    ne = 2 if strips >= 2 else 1
    waves = ne * we + (strips - ne) * wi
    assert waves <= max(slots, strips)
    cover = np.zeros((strips, ny), dtype=np.int64)
    for s, y0, y1 in chunks:
        assert 0 <= y0 < y1 <= ny
        cover[s, y0:y1] += 1
    assert (cover == 1).all()
    # a plain-path chunk needs the rows it freezes and loads inside the grid
    assert he >= R * kf + reg_period(R, ring) + ring - 2


# ---------------------------------------------------------------- 3-D register-streaming family

def reg3d_split(nz, ny, nx, ox, by, slots):
    lib = _dbg()
    lib.cs_debug_reg3d_split.restype = ctypes.c_int32
    lib.cs_debug_reg3d_split.argtypes = [ctypes.c_int64] * 3 + [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                                                ctypes.POINTER(ctypes.c_int64)]
    split = (ctypes.c_int64 * 4)()
    assert lib.cs_debug_reg3d_split(nz, ny, nx, ox, by, slots, split) == 0
    return [int(s) for s in split]


def emulate_3d_reg(v, layout, dense, x, slots):
    """k3d_reg: wave w takes y-tile w mod y-tiles, then strip, then z-chunk; it owns BY rows
    by 64 PX columns (lane-major: the lanes' runs make one flat row, extended by the column
    on either side that lanes 0 and 63 load) and streams input planes zs = z0 - 1 ..,
    loaded RING - 1 iterations ahead; plane z adds its dz = +1 taps to output plane z - 1
    (then stored), its dz = 0 taps to plane z and starts plane z + 1 with its dz = -1 taps,
    in (dy, dx) ascending order per plane.  Loads outside the array read zero on the checked
    path; the plain path must stay inside the interior's one-cell neighbourhood.  Checks
    every load and store against the array and counts the stores per cell."""
    PX, BY, RING = v["px"], v["by"], v["st"]
    OX = 64 * PX
    PER = RING if RING % 3 == 0 else 3 * RING
    box = v["shape"] == 0
    assert v["mf"] == 9 and v["R"] == 1 and v["kf"] == 1 and v["bx"] == OX and v["lds"] == 0
    ez, ey, ex = layout.ext
    oz, oy, ox = layout.org
    nz, ny, nx = layout.n
    X = x.reshape(ez, ey, ex)
    w = np.asarray(dense, dtype=np.float64).reshape(3, 3, 3)
    out = np.zeros((ez, ey, ex))
    hits = np.zeros((ez, ey, ex), dtype=np.int64)
    strips, ytiles, zchunks, zc = reg3d_split(nz, ny, nx, OX, BY, slots)
    assert strips * ytiles * zchunks <= max(slots, strips * ytiles)
    plain = 0
    for wid in range(strips * ytiles * zchunks):
        yt, rest = wid % ytiles, wid // ytiles
        strip, zch = rest % strips, rest // strips
        z0 = zch * zc
        if z0 >= nz:
            continue
        z1 = min(z0 + zc, nz)
        y0, x0 = yt * BY, strip * OX
        fast = x0 >= 1 and x0 + OX + 1 <= nx and y0 >= 1 and y0 + BY + 1 <= ny and z0 >= 1 and z1 + PER + RING <= nz
        plain += fast
        ecols = x0 - 1 + np.arange(OX + 2)  # columns x0 - 1 .. x0 + OX
        erows = y0 - 1 + np.arange(BY + 2)

        def load(z):
            if fast:
                assert 0 <= z < nz and erows[0] >= 0 and erows[-1] < ny and ecols[0] >= 0 and ecols[-1] < nx
                return X[oz + z][np.ix_(oy + erows, ox + ecols)].copy()
            ok = ((oy + erows >= 0) & (oy + erows < ey))[:, None] & ((ox + ecols >= 0) & (ox + ecols < ex))[None, :]
            if not 0 <= oz + z < ez:
                ok[:] = False
            ry = np.clip(oy + erows, 0, ey - 1)
            cx = np.clip(ox + ecols, 0, ex - 1)
            return np.where(ok, X[min(max(oz + z, 0), ez - 1)][np.ix_(ry, cx)], 0.0)

        zs, iters = z0 - 1, (z1 - z0) + 2
        acc = [np.zeros((BY, OX)) for _ in range(3)]
        ring = [load(zs + k) for k in range(RING - 1)] + [None]
        colout = x0 + np.arange(OX) < nx
        for it in range(0, iters, PER):
            for K in range(PER):
                z = zs + it + K
                ring[(K + RING - 1) % RING] = load(z + RING - 1)
                qe = ring[K % RING]
                for dz in (1, 0, -1):
                    sl = (K - dz) % 3
                    t = None if dz == -1 else acc[sl]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            if not box and (dz != 0) + (dy != 0) + (dx != 0) > 1:
                                continue
                            term = w[dz + 1, dy + 1, dx + 1] * qe[1 + dy:1 + dy + BY, 1 + dx:1 + dx + OX]
                            t = term if t is None else t + term
                    acc[sl] = t
                zo = z - 1
                if z0 <= zo < z1:
                    A = acc[(K - 1) % 3]
                    for i in range(BY):
                        if y0 + i >= ny:
                            continue
                        c = ox + x0 + np.arange(OX)[colout]
                        assert 0 <= oz + zo < ez and 0 <= oy + y0 + i < ey and (c < ex).all()
                        out[oz + zo, oy + y0 + i, c] = A[i, colout]
                        hits[oz + zo, oy + y0 + i, c] += 1
    inner = hits[oz:oz + nz, oy:oy + ny, ox:ox + nx]
    assert (inner == 1).all(), (int(inner.min()), int(inner.max()))
    assert hits.sum() == inner.sum()
    emulate_3d_reg.plain_waves = plain
    return out.reshape(-1)


VARIANTS_REG3D = [v for v in variants(9) if v["kf"] == 1]


@pytest.mark.parametrize("v", VARIANTS_REG3D, ids=[v["name"] for v in VARIANTS_REG3D])
def test_variant_emulation_3d_reg(v):
    # three strips, y-tiles and z-chunks (a plain-path wave in the middle), halos of odd
    # widths (masks at the array's edges), one ragged strip and tile, and a tiny grid
    OX, BY = v["bx"], v["by"]
    rng = np.random.default_rng(v["id"])
    w = rng.uniform(0.05, 1.0, (3, 3, 3))
    zz, yy, xx = np.meshgrid(*[np.arange(-1, 2)] * 3, indexing="ij")
    if v["shape"] == 1:
        w[((zz != 0).astype(int) + (yy != 0) + (xx != 0)) > 1] = 0.0
    w /= w.sum()
    dense = w.reshape(-1)
    ny = 3 * BY - 1
    cases = [(Layout.with_halo([40, ny, 3 * OX - 5], 1), 3 * 3 * 3, True),
             (Layout(3, [29, ny + 3, 2 * OX + 9], [2, 1, 3], [25, ny, 2 * OX + 3]), 4, False),
             (Layout.with_halo([3, 1, 5], 1), 64, False)]
    for layout, slots, want_plain in cases:
        x = oracle.fill_data(layout, v["id"] + slots)
        got = emulate_3d_reg(v, layout, dense, x, slots)
        if want_plain:
            assert emulate_3d_reg.plain_waves > 0
        offs, ww = oracle.dense_to_taps(3, 1, dense)
        want = oracle.run(layout, offs, ww, x, 1)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], slots, err)


@pytest.mark.parametrize("nz,ny,nx,ox,by,slots", [(512, 512, 512, 128, 2, 2048), (512, 512, 512, 256, 2, 1024),
                                                  (3, 1, 5, 128, 4, 64), (100, 30, 1000, 128, 2, 100000),
                                                  (64, 7, 129, 256, 4, 1)])
def test_reg3d_split_covers_once(nz, ny, nx, ox, by, slots):
    strips, ytiles, zchunks, zc = reg3d_split(nz, ny, nx, ox, by, slots)
    assert strips * ox >= nx > (strips - 1) * ox and ytiles * by >= ny > (ytiles - 1) * by
    assert zchunks * zc >= nz > (zchunks - 1) * zc
    assert zc >= min(8, nz) and strips * ytiles * zchunks <= max(slots, strips * ytiles)


def emulate_3f_reg(v, layout, dense, x, slots):
    """k3f_reg (KF fused steps, zero halo): wave w takes y-tile w mod y-tiles, then strip,
    then z-chunk; it owns BY output rows and a strip of 64 PX columns starting KF columns
    left of its outputs (lane-major flat row; the lane shifts bring zero past the wave's
    ends), streams input planes z0 - KF .., loaded RING - 1 iterations ahead, rows y0 - KF
    .. y0 + BY + KF - 1; level l (rows y0 - KF + l .. y0 + BY + KF - l - 1) takes its input
    plane p = z - (l - 1) into its planes p + 1, p, p - 1, and its plane z - l, then
    complete, is masked to zero outside the interior on the checked path and feeds level
    l + 1; level KF's plane z - KF is stored on the output columns (the plain path in pairs
    of a lane's cells when PX and KF are even).  Checks loads and stores against the array,
    counts the stores per cell, and that the plain path's intermediate cells are interior."""
    PX, BY, RING, KF = v["px"], v["by"], v["st"], v["kf"]
    OX = 64 * PX - 2 * KF
    NR = BY + 2 * KF
    PER = RING if RING % 3 == 0 else 3 * RING
    box = v["shape"] == 0
    assert v["mf"] == 9 and v["R"] == 1 and KF >= 2 and v["bx"] == OX and v["lds"] == 0
    ez, ey, ex = layout.ext
    oz, oy, ox = layout.org
    nz, ny, nx = layout.n
    X = x.reshape(ez, ey, ex)
    w = np.asarray(dense, dtype=np.float64).reshape(3, 3, 3)
    out = np.zeros((ez, ey, ex))
    hits = np.zeros((ez, ey, ex), dtype=np.int64)
    strips, ytiles, zchunks, zc = reg3d_split(nz, ny, nx, OX, BY, slots)
    plain = 0
    W = 64 * PX
    d = np.arange(W)

    def rows_of(lv):
        return BY + 2 * (KF - lv)

    def shifted(q):  # q[rows][W] -> extended rows [rows][W + 2] with zero past the wave's ends
        return np.concatenate([np.zeros((q.shape[0], 1)), q, np.zeros((q.shape[0], 1))], axis=1)

    def add_taps(src, T, dz, start, nout):
        t = None if start else T
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if not box and (dz != 0) + (dy != 0) + (dx != 0) > 1:
                    continue
                term = w[dz + 1, dy + 1, dx + 1] * src[1 + dy:1 + dy + nout, 1 + dx:1 + dx + W]
                t = term if t is None else t + term
        return t

    for wid in range(strips * ytiles * zchunks):
        yt, rest = wid % ytiles, wid // ytiles
        strip, zch = rest % strips, rest // strips
        z0 = zch * zc
        if z0 >= nz:
            continue
        z1 = min(z0 + zc, nz)
        y0, xs = yt * BY, strip * OX - KF
        fast = (xs >= 0 and xs + W <= nx and y0 >= KF and y0 + BY + KF <= ny and z0 >= KF
                and z1 + KF + PER + RING - 2 <= nz)
        plain += fast
        cols = xs + d
        rows = y0 - KF + np.arange(NR)
        colin = (cols >= 0) & (cols < nx)
        colout = (d >= KF) & (d < KF + OX) & (cols < nx)

        def load(z):
            if fast:
                assert 0 <= z < nz and rows[0] >= 0 and rows[-1] < ny and cols[0] >= 0 and cols[-1] < nx
                return X[oz + z][np.ix_(oy + rows, ox + cols)].copy()
            cz = min(max(oz + z, 0), ez - 1)
            return X[cz][np.ix_(np.clip(oy + rows, 0, ey - 1), np.clip(ox + cols, 0, ex - 1))]

        zs, iters = z0 - KF, (z1 - z0) + 2 * KF
        acc = [[np.zeros((rows_of(lv), W)) for _ in range(3)] for lv in range(1, KF + 1)]
        ring = [load(zs + k) for k in range(RING - 1)] + [None]
        for it in range(0, iters, PER):
            for K in range(PER):
                z = zs + it + K
                ring[(K + RING - 1) % RING] = load(z + RING - 1)
                src = ring[K % RING]
                for lv in range(1, KF + 1):
                    A = acc[lv - 1]
                    se = shifted(src)
                    for dz in (1, 0, -1):
                        sl = (K - (lv - 1) - dz) % 3
                        A[sl] = add_taps(se, A[sl], dz, dz == -1, rows_of(lv))
                    P = A[(K - lv) % 3]
                    if lv < KF:
                        p = z - lv
                        ry = y0 - (KF - lv) + np.arange(rows_of(lv))
                        if fast:
                            if z0 - (KF - lv) <= p <= z1 + (KF - lv) - 1:  # planes the outputs use
                                assert 0 <= p < nz and ry[0] >= 0 and ry[-1] < ny and colin.all()
                        else:
                            rin = (ry >= 0) & (ry < ny) & (0 <= p < nz)
                            P = np.where(rin[:, None] & colin[None, :], P, 0.0)
                            A[(K - lv) % 3] = P
                    src = P
                zo = z - KF
                if z0 <= zo < z1:
                    Aout = acc[KF - 1][(K - KF) % 3]
                    st = colout.copy()
                    if fast and PX % 2 == 0 and KF % 2 == 0:
                        st = np.repeat(colout[0::2], 2)
                    for i in range(BY):
                        if y0 + i >= ny:
                            continue
                        c = ox + cols[st]
                        assert 0 <= oz + zo < ez and 0 <= oy + y0 + i < ey and (c >= 0).all() and (c < ex).all()
                        out[oz + zo, oy + y0 + i, c] = Aout[i, st]
                        hits[oz + zo, oy + y0 + i, c] += 1
    inner = hits[oz:oz + nz, oy:oy + ny, ox:ox + nx]
    assert (inner == 1).all(), (int(inner.min()), int(inner.max()))
    assert hits.sum() == inner.sum()
    emulate_3f_reg.plain_waves = plain
    return out.reshape(-1)


VARIANTS_REG3F = [v for v in variants(9) if v["kf"] > 1]


@pytest.mark.parametrize("v", VARIANTS_REG3F, ids=[v["name"] for v in VARIANTS_REG3F])
def test_variant_emulation_3f_reg(v):
    # zero halo (the plans fuse only then): a plain-path wave in the middle of the large
    # grid, halos of odd widths, ragged strips and tiles, and tiny grids
    OX, BY = v["bx"], v["by"]
    rng = np.random.default_rng(v["id"])
    w = rng.uniform(0.05, 1.0, (3, 3, 3))
    zz, yy, xx = np.meshgrid(*[np.arange(-1, 2)] * 3, indexing="ij")
    if v["shape"] == 1:
        w[((zz != 0).astype(int) + (yy != 0) + (xx != 0)) > 1] = 0.0
    w /= w.sum()
    dense = w.reshape(-1)
    KF = v["kf"]
    ny = 3 * BY + 2 * KF
    three_chunks = 3 * ((ny + BY - 1) // BY) * 3  # strips x y-tiles x 3 z-chunks
    cases = [(Layout.with_halo([48, ny, 3 * OX - 5], 1), three_chunks, True),
             (Layout(3, [29, ny + 3, 2 * OX + 9], [2, 1, 3], [25, ny, 2 * OX + 3]), 4, False),
             (Layout.with_halo([3, 1, 5], 1), 64, False), (Layout.with_halo([1, 2, 130], 2), 8, False)]
    for layout, slots, want_plain in cases:
        x = np.where(layout.interior_mask(), oracle.fill_data(layout, v["id"] + slots), 0.0)
        got = emulate_3f_reg(v, layout, dense, x, slots)
        if want_plain:
            assert emulate_3f_reg.plain_waves > 0
        offs, ww = oracle.dense_to_taps(3, 1, dense)
        want = oracle.run(layout, offs, ww, x, KF)
        gi, wi = layout.interior(got), layout.interior(want)
        err = np.max(np.abs(gi - wi)) / np.max(np.abs(wi))
        assert err < 1e-13, (v["name"], slots, err)
