"""Independent numpy restatements used to pin the C oracle (tests only).

`run_numpy` is written from the reference loops' definition (not from
cs_oracle.c): every interior point is w0*x0 + w1*x1 + ... summed left to right in
row-major offset order, halo cells are never written, ping-pong buffers with the
second buffer starting as zeros (src/2d/main.cu:57-115, :381-393).
"""
import numpy as np


def run_numpy(host_in, layout, offs, w, steps, second=None):
    e = layout.ext
    o = layout.org
    n = layout.n
    bufs = [np.array(host_in, dtype=np.float64).reshape(e),
            np.zeros(e) if second is None else np.array(second, dtype=np.float64).reshape(e)]
    z0, y0, x0 = o
    for t in range(steps):
        src, dst = bufs[t % 2], bufs[(t + 1) % 2]
        acc = None
        for (dz, dy, dx), wt in zip(offs, w):
            sl = src[z0 + dz:z0 + dz + n[0], y0 + dy:y0 + dy + n[1], x0 + dx:x0 + dx + n[2]]
            term = wt * sl
            acc = term if acc is None else acc + term
        dst[z0:z0 + n[0], y0:y0 + n[1], x0:x0 + n[2]] = acc
    return bufs[steps % 2].reshape(-1)


def brute_point(a3, z, y, x, offs, w):
    """Single point, scalar Python loop (tiny cases only)."""
    acc = None
    for (dz, dy, dx), wt in zip(offs, w):
        term = float(wt) * float(a3[z + dz, y + dy, x + dx])
        acc = term if acc is None else acc + term
    return acc


def splitmix64_py(state, count):
    """Pure-Python SplitMix64 (Steele, Lea & Flood 2014): returns (outputs, state)."""
    M = (1 << 64) - 1
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out, state


def rel_errors(g, c):
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    c = np.asarray(c, dtype=np.float64).reshape(-1)
    d = np.abs(g - c)
    mx = float(np.max(np.abs(c))) if c.size else 0.0
    rel_max = float(np.max(d)) / mx if mx > 0 else float(np.max(d, initial=0.0))
    l2c = float(np.linalg.norm(c))
    rel_l2 = float(np.linalg.norm(g - c)) / l2c if l2c > 0 else float(np.linalg.norm(g - c))
    return rel_max, rel_l2
