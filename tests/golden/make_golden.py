"""Writes the golden fixtures in tests/golden/*.npz (run from the repo root:
python tests/golden/make_golden.py).

Inputs come from a pure-Python SplitMix64 (tests/refimpl.py), outputs from the
numpy restatement of the reference loops (tests/refimpl.run_numpy) -- neither
uses the C oracle, so the fixtures pin it.  The reference repository has no
tests or golden files of its own (its CHECK_ERROR block is disabled and does not
compile, src/2d/main.cu:23,375-406).

Cases cover every BASELINE shape on small odd sizes, plus the reference's own
2-D padded layout (rows m+6, cols n+8, interior at (3, 4)) with a non-zero halo,
where the halo alternates between the input's values and zeros from step to
step exactly as the reference's two device buffers do (src/2d/gpu.cu:273-277).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from convstencil_amd.shapes import Layout, shape_offsets, SHAPES  # noqa: E402
from refimpl import run_numpy, splitmix64_py  # noqa: E402

M64 = (1 << 64) - 1


def stream_state(seed, stream):
    outs, _ = splitmix64_py((seed * 2 + stream) & M64, 1)
    return outs[0]


def data(layout, seed, halo_random=False):
    a = np.zeros(layout.ext)
    st = stream_state(seed, 0)
    n = layout.n
    vals, st = splitmix64_py(st, n[0] * n[1] * n[2])
    inter = np.array([(v >> 11) * 2.0 ** -53 * 2.0 - 1.0 for v in vals]).reshape(n)
    o = layout.org
    if halo_random:
        hv, _ = splitmix64_py(st, layout.cells)
        a = np.array([float((v >> 11) % 100) for v in hv]).reshape(layout.ext)
    a[o[0]:o[0] + n[0], o[1]:o[1] + n[1], o[2]:o[2] + n[2]] = inter
    return a.reshape(-1)


def weights(count, seed):
    st = stream_state(seed, 1)
    vals, _ = splitmix64_py(st, count)
    w = [((v >> 11) + 0.5) * 2.0 ** -53 for v in vals]
    s = 0.0
    for x in w:
        s = s + x
    return np.array([x / s for x in w])


CASES = [
    ("1d1r", [37], 4, 11),
    ("1d2r", [41], 3, 12),
    ("box2d1r", [13, 19], 3, 13),
    ("star2d1r", [10, 21], 3, 14),
    ("star2d3r", [11, 17], 2, 15),
    ("box2d3r", [9, 12], 2, 16),
    ("star3d1r", [5, 6, 7], 2, 17),
    ("box3d1r", [4, 5, 6], 2, 18),
]


def main():
    for name, sizes, steps, seed in CASES:
        ndim, radius, kind = SHAPES[name]
        L = Layout.with_halo(sizes, radius)
        offs = shape_offsets(name)
        w = weights(len(offs), seed)
        x = data(L, seed)
        y = run_numpy(x, L, offs, w, steps)
        np.savez(os.path.join(HERE, f"{name}.npz"), ext=np.array(L.ext), org=np.array(L.org), n=np.array(L.n),
                 offs=np.array(offs, dtype=np.int32), w=w, steps=np.array(steps), seed=np.array(seed), x=x, y=y)
    # reference 2-D layout, 49 weights, random non-zero halo (alternating-halo semantics)
    m, n = 7, 10
    L = Layout(2, [m + 6, n + 8], [3, 4], [m, n])
    offs = [(0, dy, dx) for dy in range(-3, 4) for dx in range(-3, 4)]
    w = weights(49, 21)
    x = data(L, 21, halo_random=True)
    y = run_numpy(x, L, offs, w, 3)
    np.savez(os.path.join(HERE, "ref2d_layout_randhalo.npz"), ext=np.array(L.ext), org=np.array(L.org),
             n=np.array(L.n), offs=np.array(offs, dtype=np.int32), w=w, steps=np.array(3), seed=np.array(21), x=x, y=y)
    # first SplitMix64 outputs of the data and weight streams for seed 1234
    st_d, st_w = stream_state(1234, 0), stream_state(1234, 1)
    np.savez(os.path.join(HERE, "splitmix64.npz"),
             data=np.array(splitmix64_py(st_d, 8)[0], dtype=np.uint64),
             weights=np.array(splitmix64_py(st_w, 8)[0], dtype=np.uint64))
    print("golden fixtures written")


if __name__ == "__main__":
    main()
