// cs_ktable.h -- stencil2row k-step tables shared by host and device code.
//
// Both FP64 MFMA shapes on gfx950 compute D += A * B with the input tile as A
// and a banded Toeplitz slice of the weights as B (registers, whole kernel).
// The reference builds the same kind of Toeplitz for m8n8k4 as two half
// matrices ("dual tessellation", src/2d/gpu.cu:246-266) and scatters each input
// element into two stencil2row copies in shared memory (src/2d/gpu.cu:287-310);
// here A's rows are rows of the LDS tile as loaded, so there is no scatter.
//
// v_mfma_f64_16x16x4 (lane l: A[l%16][l/16], B[l/16][l%16], D[4v + l/16][l%16]):
//   a 16x16 output block; each k-step is one stencil row dy and four consecutive
//   input columns c0..c0+3 relative to the block's first output column:
//     A[i][kk] = tile[y + i + dy][x + c0 + kk],  B[kk][n] = W[dy][c0 + kk - n].
//   A stencil row with taps dx in [lo, hi] needs columns [lo, 15 + hi], i.e.
//   ceil((16 + hi - lo) / 4) k-steps (make_ktable).
//
// v_mfma_f64_4x4x4 (4 blocks; lane l: block (l>>2)&3, A[l&3][l>>4],
//   B[l>>4][l&3], D[l>>4][l&3]): four 4x4 output blocks side by side form a
//   4-row x 16-column strip.  K is the flattened list of (dy, c) pairs a 4-wide
//   output block needs -- c in [lo, 3 + hi] for every stencil row -- padded to a
//   multiple of 4 with zero-weight entries (make_ftable).  Per lane, k-step ks
//   reads tile[y + i + dy_f][x + 4b + c_f] with f = 4 ks + (l >> 4), and
//   B[f][j] = W[dy_f][c_f - j].  For box2d1r this is 5 k-steps per 64 outputs
//   (40 executed flops per point-step against 18 useful), against 15 k-steps per
//   256 outputs (120 flops) with the 16x16x4 shape.
#pragma once

namespace cs {

enum Shape2D : int { SHAPE_BOX = 0, SHAPE_STAR = 1 };

struct KTable {
  int n;        // number of k-steps (16x16x4) or of flattened (dy, c) entries (4x4x4, padded)
  int dy[128];  // stencil row offset in [-R, R]
  int c0[128];  // input column relative to the output block's first column
  int live[128];// 1 for a real entry, 0 for padding (4x4x4 tables only)
};

// Taps present in row dy of a stencil of this shape: [lo, hi], or lo > hi if none.
constexpr void row_span(int R, int shape, int dy, int &lo, int &hi) {
  if (shape == SHAPE_BOX || dy == 0) {
    lo = -R;
    hi = R;
  } else {
    lo = 0;
    hi = 0;
  }
}

constexpr KTable make_ktable(int R, int shape) {
  KTable t{};
  t.n = 0;
  for (int dy = -R; dy <= R; ++dy) {
    int lo = 0, hi = -1;
    row_span(R, shape, dy, lo, hi);
    if (lo > hi) continue;
    const int chunks = (16 + hi - lo + 3) / 4;
    for (int q = 0; q < chunks; ++q) {
      t.dy[t.n] = dy;
      t.c0[t.n] = lo + 4 * q;
      t.live[t.n] = 1;
      ++t.n;
    }
  }
  return t;
}

constexpr KTable make_ftable(int R, int shape) {
  KTable t{};
  t.n = 0;
  for (int dy = -R; dy <= R; ++dy) {
    int lo = 0, hi = -1;
    row_span(R, shape, dy, lo, hi);
    if (lo > hi) continue;
    for (int c = lo; c <= 3 + hi; ++c) {
      t.dy[t.n] = dy;
      t.c0[t.n] = c;
      t.live[t.n] = 1;
      ++t.n;
    }
  }
  while (t.n % 4 != 0) {  // padding entries read the block's own first cell with weight 0
    t.dy[t.n] = 0;
    t.c0[t.n] = 0;
    t.live[t.n] = 0;
    ++t.n;
  }
  return t;
}

// Largest column read past the block's last output column (x + 15) by a 16x16x4 table.
constexpr int ktable_overrun(const KTable &t) {
  int m = 0;
  for (int s = 0; s < t.n; ++s) {
    int last = t.c0[s] + 3 - 15;
    if (last > m) m = last;
  }
  return m;
}

}  // namespace cs

namespace cs {

// ---- 3-D tables (v_mfma_f64_4x4x4, 4 x 16 output strips of one z-plane) ----
//
// K is grouped by dz so that every k-step reads a single LDS plane (the plane's
// ring slot is then uniform over the wave); inside a dz group the (dy, c) entries
// are flattened as in make_ftable and padded to a multiple of 4.  Shape classes:
// box (every tap), star (taps on the axes), diamond (|dz| + |dy| + |dx| <= R: the
// reference's 63-point star3d1r is the radius-3 diamond, src/3d/main.cu:42-113).
enum Shape3D : int { SHAPE3_BOX = 0, SHAPE3_STAR = 1, SHAPE3_DIAMOND = 2 };

struct KTable3D {
  int n;             // k-steps
  int dz[160];       // plane of each k-step
  int dy[160][4];    // per k-step, per k lane-group: row offset
  int c[160][4];     // column offset relative to the 4-wide block
  int live[160][4];  // 0 for padding entries (zero weight)
};

constexpr void row_span_3d(int R, int shape, int dz, int dy, int &lo, int &hi) {
  const int adz = dz < 0 ? -dz : dz, ady = dy < 0 ? -dy : dy;
  lo = 0;
  hi = -1;
  if (shape == SHAPE3_BOX) {
    lo = -R;
    hi = R;
  } else if (shape == SHAPE3_STAR) {
    if (dz == 0 && dy == 0) {
      lo = -R;
      hi = R;
    } else if (dz == 0 || dy == 0) {
      lo = 0;
      hi = 0;
    }
  } else {
    const int rem = R - adz - ady;
    if (rem >= 0) {
      lo = -rem;
      hi = rem;
    }
  }
}

constexpr KTable3D make_ktable3d(int R, int shape) {
  KTable3D t{};
  t.n = 0;
  for (int dz = -R; dz <= R; ++dz) {
    int fdy[128] = {}, fc[128] = {};
    int m = 0;
    for (int dy = -R; dy <= R; ++dy) {
      int lo = 0, hi = -1;
      row_span_3d(R, shape, dz, dy, lo, hi);
      for (int c = lo; c <= 3 + hi && lo <= hi; ++c) {
        fdy[m] = dy;
        fc[m] = c;
        ++m;
      }
    }
    for (int q = 0; q < (m + 3) / 4; ++q) {
      t.dz[t.n] = dz;
      for (int k = 0; k < 4; ++k) {
        const int f = 4 * q + k;
        t.dy[t.n][k] = f < m ? fdy[f] : 0;
        t.c[t.n][k] = f < m ? fc[f] : 0;
        t.live[t.n][k] = f < m ? 1 : 0;
      }
      ++t.n;
    }
  }
  return t;
}

}  // namespace cs
