"""Stencil shapes, layouts and weight helpers (host side).

Shape names are the reference's CLI names (src/1d/main.cu:73-81,
src/2d/main.cu:139-160, src/3d/main.cu:494-502).  In the reference the `*1r`
shapes run pre-fused radius-3 weights; here a shape is the physical stencil
(1d1r = 3 points, box2d1r = 9 points, ...) and fusion happens in the kernels.
Taps are listed in row-major offset order (z outer, then y, then x), which is
also the order in which the CPU oracle sums them.
"""
import numpy as np

# name -> (ndim, radius, kind); kind "box" = full cube, "star" = axes only
SHAPES = {
    "1d1r": (1, 1, "box"),
    "1d2r": (1, 2, "box"),
    "star2d1r": (2, 1, "star"),
    "box2d1r": (2, 1, "box"),
    "star2d3r": (2, 3, "star"),
    "box2d3r": (2, 3, "box"),
    "star3d1r": (3, 1, "star"),
    "box3d1r": (3, 1, "box"),
}


def offsets(ndim, radius, kind):
    """Tap offsets (dz, dy, dx) in row-major order."""
    r = range(-radius, radius + 1)
    zs = r if ndim == 3 else [0]
    ys = r if ndim >= 2 else [0]
    out = []
    for dz in zs:
        for dy in ys:
            for dx in r:
                nz = (dz != 0) + (dy != 0) + (dx != 0)
                if kind == "star" and nz > 1:
                    continue
                out.append((dz, dy, dx))
    return out


def shape_offsets(name):
    ndim, radius, kind = SHAPES[name]
    return offsets(ndim, radius, kind)


def dense_weights(ndim, radius, offs, w):
    """Dense (2r+1)^ndim weight cube (row-major z, y, x) from a tap list."""
    d = 2 * radius + 1
    cube = np.zeros((d,) * ndim, dtype=np.float64)
    for (dz, dy, dx), v in zip(offs, w):
        idx = {1: (dx + radius,), 2: (dy + radius, dx + radius), 3: (dz + radius, dy + radius, dx + radius)}[ndim]
        cube[idx] = v
    return cube.reshape(-1)


class Layout:
    """Padded array layout: ext/org/n slowest first, padded to 3 entries."""

    def __init__(self, ndim, ext, org, n):
        self.ndim = int(ndim)
        pad = 3 - self.ndim
        self.ext = tuple([1] * pad + [int(v) for v in ext])
        self.org = tuple([0] * pad + [int(v) for v in org])
        self.n = tuple([1] * pad + [int(v) for v in n])

    @classmethod
    def with_halo(cls, sizes, halo, extra_x=0):
        """Interior `sizes` (slowest first) with a halo of `halo` cells on every side
        and `extra_x` more columns at the end of each row."""
        sizes = [int(s) for s in sizes]
        ext = [s + 2 * halo for s in sizes]
        ext[-1] += extra_x
        return cls(len(sizes), ext, [halo] * len(sizes), sizes)

    @property
    def cells(self):
        return self.ext[0] * self.ext[1] * self.ext[2]

    @property
    def points(self):
        return self.n[0] * self.n[1] * self.n[2]

    def shape3(self):
        return self.ext

    def interior(self, a):
        """View of the interior of a flat padded array."""
        v = np.asarray(a).reshape(self.ext)
        return v[self.org[0]:self.org[0] + self.n[0], self.org[1]:self.org[1] + self.n[1],
                 self.org[2]:self.org[2] + self.n[2]]

    def interior_mask(self):
        """Flat boolean mask of the interior cells of a padded array."""
        m = np.zeros(self.ext, dtype=bool)
        m[self.org[0]:self.org[0] + self.n[0], self.org[1]:self.org[1] + self.n[1],
          self.org[2]:self.org[2] + self.n[2]] = True
        return m.reshape(-1)

    def to_c(self):
        from . import _lib
        L = _lib.cs_layout()
        L.ndim = self.ndim
        for d in range(3):
            L.ext[d], L.org[d], L.n[d] = self.ext[d], self.org[d], self.n[d]
        return L

    def __repr__(self):
        return f"Layout(ndim={self.ndim}, ext={self.ext}, org={self.org}, n={self.n})"


# ---- the reference's weight expansions for its pre-fused CLI shapes ----

def expand_1d1r(v):
    """3 weights -> the 7 coefficients of v*v*v, as src/1d/main.cu:133-140 writes them."""
    v0, v1, v2 = (float(x) for x in v)
    return np.array([
        v0 * v0 * v0,
        3 * v0 * v0 * v1,
        3 * v0 * v0 * v2 + 3 * v0 * v1 * v1,
        6 * v0 * v1 * v2 + v1 * v1 * v1,
        3 * v0 * v2 * v2 + 3 * v1 * v1 * v2,
        3 * v1 * v2 * v2,
        v2 * v2 * v2,
    ])


def expand_1d2r(v):
    """5 weights -> the 9 coefficients of v*v, as src/1d/main.cu:141-151 writes them."""
    v0, v1, v2, v3, v4 = (float(x) for x in v)
    return np.array([
        v0 * v0,
        2 * v0 * v1,
        2 * v0 * v2 + v1 * v1,
        2 * v0 * v3 + 2 * v1 * v2,
        2 * v0 * v4 + 2 * v1 * v3 + v2 * v2,
        2 * v1 * v4 + 2 * v2 * v3,
        2 * v2 * v4 + v3 * v3,
        2 * v3 * v4,
        v4 * v4,
    ])


def compose(kernel, times, ndim):
    """Exact `times`-fold self-convolution of a (2r+1)^ndim weight cube."""
    k = np.asarray(kernel, dtype=np.float64)
    out = k.copy()
    for _ in range(times - 1):
        out = _full_convolve(out, k)
    return out


def _full_convolve(a, b):
    shape = tuple(sa + sb - 1 for sa, sb in zip(a.shape, b.shape))
    out = np.zeros(shape)
    for idx in np.ndindex(*b.shape):
        sl = tuple(slice(i, i + sa) for i, sa in zip(idx, a.shape))
        out[sl] += b[idx] * a
    return out
