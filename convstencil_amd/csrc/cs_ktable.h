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

namespace cs {

// ---- dilated 4x4x4 table (2-D FP64, 8 x 8 output blocks) ----
//
// One v_mfma_f64_4x4x4 (4 blocks b, D_b[m][n] = sum_k A_b[m][k] B_b[k][n]) computes
// an 8 x 8 output block: output (b, m, n) is cell (y + 2b + (n >> 1), x + 2m + (n & 1)),
// i.e. the point (y + 2b, x + 2m) moved by v_n, v_n in {0, 1}^2.  With K running over
// the dilated footprint T = {d + v : d a tap, v in {0, 1}^2} (tap d, t = d + v):
//   A_b[m][t] = tile[y + 2b + t.y][x + 2m + t.x],   B[t][n] = W[t - v_n] (0 off the stencil),
// B being the same for the four blocks.  Lane l holds A_b[m][k] with k = l >> 4,
// b = (l >> 2) & 3, m = l & 3, so the 16 lanes of one k read four 64-byte row segments
// two rows apart (the pitch is chosen with the LDS model below), and
// B[k][n] with n = l & 3; D_b[m][n] is in lane 16 m + 4 b + n.
//
// T has (2R + 2)^2 entries for a box (16 for box2d1r: 4 k-steps per 64 outputs,
// 16 executed MACs per point-step against 9 useful), 8R + 4 for a star.  Entries
// are grouped into 16-byte pairs -- t and t + (0, 1), read with one ds_read_b128 and
// fed to two k-steps -- which must start at an even LDS address: every pair starts
// at a column t.x of parity pp, and blocks start at columns x with x + CS + pp even
// (CS: the tile's storage column shift).  A row of T whose ends have the wrong parity
// leaves single entries (8-byte reads), or a pair padded with a zero-weight column;
// the split/pad choice minimises k-steps, then LDS reads.  Pair steps (two k-steps
// each) come first, single k-steps last, four entries (one per k lane group) per
// step; unused slots repeat a live address with zero weight (live = 0).
struct DTable {
  int map;             // lane -> cell map of the blocks (dil_row / dil_col): 0 rows, 1 skew
  int np;              // pair steps
  int ns;              // single k-steps
  int pp;              // parity of every pair's first column t.x
  int pty[16][4], ptx[16][4], plive[16][4];
  int sty[16][4], stx[16][4], slive[16][4];
  int tmin, tmax;      // rows t.y read, relative to the block row 2b
  int cmin, cmax;      // columns t.x read (a pair's second column included), relative to 2m
};

// Where the lanes of a v_mfma_f64_4x4x4 instruction sit in an 8 x 8 output block.  Lane
// l holds A_b[l & 3][l >> 4] and D_b[l >> 4][l & 3] of block b = (l >> 2) & 3; the pair
// (block b, A/D row i) owns one 2 x 2 sub-block, at row pair dil_row(b, i) and column pair
// dil_col(b, i) (any bijection of (b, i) onto the 4 x 4 sub-blocks computes the same
// stencil: the B fragments, the same for every block, only depend on the 2 x 2 offset of
// D's column index).  Map 0 puts block b on row pair b and row i on column pair i (the
// 3-D kernels); map 1 puts row i on row pair i and skews the column pair by the block,
// (b + i) mod 4, which together with the order of each step's taps over the four lane
// groups (order_dtable) makes both the 16-byte A reads and the 8-byte result writes free
// of LDS bank conflicts at pitches of 8 (mod 16) doubles (lds_cost below).
constexpr int dil_row(int map, int b, int i) { return map == 1 ? i : b; }
constexpr int dil_col(int map, int b, int i) { return map == 1 ? ((b + i) & 3) : i; }

// Row spans of a 2-D tap set of radius R: taps (dy, dx) with dx in [lo[dy + R], hi[dy + R]]
// (lo > hi: no taps in that row).  The dilated tables below are built from these.
struct Spans {
  int R;
  int lo[9], hi[9];
};

constexpr Spans spans_2d(int R, int shape) {
  Spans s{};
  s.R = R;
  for (int dy = -R; dy <= R; ++dy) row_span(R, shape, dy, s.lo[dy + R], s.hi[dy + R]);
  return s;
}

// Columns [lo, hi] of row ty of the dilated footprint T (lo > hi if the row is empty).
constexpr void dilated_row(const Spans &sp, int ty, int &lo, int &hi) {
  lo = 1 << 20;
  hi = -(1 << 20);
  for (int vy = 0; vy <= 1; ++vy) {
    const int dy = ty - vy;
    if (dy < -sp.R || dy > sp.R) continue;
    const int a = sp.lo[dy + sp.R], b = sp.hi[dy + sp.R];
    if (a > b) continue;
    lo = a < lo ? a : lo;
    hi = b + 1 > hi ? b + 1 : hi;
  }
}

constexpr int ceil4(int v) { return (v + 3) / 4; }

// pp_force: -1 chooses the pair parity, 0 or 1 imposes it (3-D planes share one parity).
constexpr DTable make_dtable_spans(const Spans &sp, int pp_force) {
  const int R = sp.R;
  DTable best{};
  int best_ks = 1 << 30, best_rd = 1 << 30;
  for (int pp = 0; pp < 2; ++pp) {
    if (pp_force >= 0 && pp != pp_force) continue;
    int pty[64] = {}, ptx[64] = {}, np0 = 0;
    int sty[64] = {}, stx[64] = {}, side[64] = {}, ns0 = 0;
    for (int ty = -R; ty <= R + 1; ++ty) {
      int lo = 0, hi = -1;
      dilated_row(sp, ty, lo, hi);
      if (lo > hi) continue;
      int a = lo;
      if ((((a % 2) + 2) % 2) != pp) {  // left end of the wrong parity
        sty[ns0] = ty;
        stx[ns0] = a;
        side[ns0] = -1;
        ++ns0;
        ++a;
      }
      for (; a + 1 <= hi; a += 2) {
        pty[np0] = ty;
        ptx[np0] = a;
        ++np0;
      }
      if (a == hi) {  // right end left over
        sty[ns0] = ty;
        stx[ns0] = a;
        side[ns0] = 1;
        ++ns0;
      }
    }
    // split the last j pairs into singles; pad the first k singles into pairs
    for (int j = 0; j <= np0; ++j)
      for (int k = 0; k <= ns0; ++k) {
        const int P = np0 - j + k, S = ns0 - k + 2 * j;
        const int ks = 2 * ceil4(P) + ceil4(S), rd = ceil4(P) + ceil4(S);
        if (ks > best_ks || (ks == best_ks && rd >= best_rd)) continue;
        if (P > 64 || S > 64) continue;
        best_ks = ks;
        best_rd = rd;
        DTable t{};
        t.pp = pp;
        int fy[64] = {}, fx[64] = {}, gy[64] = {}, gx[64] = {};
        int nP = 0, nS = 0;
        for (int i = 0; i < np0 - j; ++i) {
          fy[nP] = pty[i];
          fx[nP] = ptx[i];
          ++nP;
        }
        for (int i = 0; i < k; ++i) {  // a single at the left end pairs with the column before it
          fy[nP] = sty[i];
          fx[nP] = side[i] < 0 ? stx[i] - 1 : stx[i];
          ++nP;
        }
        for (int i = k; i < ns0; ++i) {
          gy[nS] = sty[i];
          gx[nS] = stx[i];
          ++nS;
        }
        for (int i = np0 - j; i < np0; ++i)
          for (int h = 0; h < 2; ++h) {
            gy[nS] = pty[i];
            gx[nS] = ptx[i] + h;
            ++nS;
          }
        t.np = ceil4(nP);
        t.ns = ceil4(nS);
        for (int e = 0; e < 4 * t.np; ++e) {
          const bool live = e < nP;
          const int src = live ? e : 4 * (e / 4);
          t.pty[e / 4][e % 4] = fy[src];
          t.ptx[e / 4][e % 4] = fx[src];
          t.plive[e / 4][e % 4] = live ? 1 : 0;
        }
        for (int e = 0; e < 4 * t.ns; ++e) {
          const bool live = e < nS;
          const int src = live ? e : 4 * (e / 4);
          t.sty[e / 4][e % 4] = gy[src];
          t.stx[e / 4][e % 4] = gx[src];
          t.slive[e / 4][e % 4] = live ? 1 : 0;
        }
        t.tmin = 1 << 20;
        t.tmax = -(1 << 20);
        t.cmin = 1 << 20;
        t.cmax = -(1 << 20);
        for (int e = 0; e < nP; ++e) {
          t.tmin = fy[e] < t.tmin ? fy[e] : t.tmin;
          t.tmax = fy[e] > t.tmax ? fy[e] : t.tmax;
          t.cmin = fx[e] < t.cmin ? fx[e] : t.cmin;
          t.cmax = fx[e] + 1 > t.cmax ? fx[e] + 1 : t.cmax;
        }
        for (int e = 0; e < nS; ++e) {
          t.tmin = gy[e] < t.tmin ? gy[e] : t.tmin;
          t.tmax = gy[e] > t.tmax ? gy[e] : t.tmax;
          t.cmin = gx[e] < t.cmin ? gx[e] : t.cmin;
          t.cmax = gx[e] > t.cmax ? gx[e] : t.cmax;
        }
        best = t;
      }
  }
  return best;
}

constexpr int lds_cost(const int (&addr)[64], int bytes, bool write);

// Model cost of the A reads of one step of table t (taps of lane group q = order[q]) at pitch
// px under t's map.
constexpr int dstep_read_cost(const DTable &t, int step, const int (&order)[4], int px) {
  const bool pair = step < t.np;
  int a[64] = {};
  for (int l = 0; l < 64; ++l) {
    const int q = order[l >> 4], b = (l >> 2) & 3, i = l & 3;
    const int ty = pair ? t.pty[step][q] : t.sty[step - t.np][q];
    const int tx = pair ? t.ptx[step][q] : t.stx[step - t.np][q];
    a[l] = ((2 * dil_row(t.map, b, i) + ty + 16) * px + 2 * dil_col(t.map, b, i) + tx + 16) * 8;
  }
  return lds_cost(a, pair ? 16 : 8, false);
}

// Puts each step's four taps on the lane groups in the order with the cheapest reads at
// the pitches the skewed map favours (8 mod 16 doubles; two samples); ties keep the order
// of construction.  Only the order within a step changes, so k-steps, reads and the
// footprint stay those of the table.
constexpr DTable order_dtable(DTable t) {
  for (int st = 0; st < t.np + t.ns; ++st) {
    const bool pair = st < t.np;
    int best[4] = {0, 1, 2, 3}, best_c = 1 << 30;
    for (int p = 0; p < 24; ++p) {
      // permutation p of {0, 1, 2, 3} (Lehmer code)
      int avail[4] = {0, 1, 2, 3}, ord[4] = {}, k = p, n = 4;
      for (int j = 0; j < 4; ++j) {
        int f = 1;
        for (int x = 2; x < n; ++x) f *= x;
        const int idx = k / f;
        k %= f;
        ord[j] = avail[idx];
        for (int x = idx; x < n - 1; ++x) avail[x] = avail[x + 1];
        --n;
      }
      const int c = dstep_read_cost(t, st, ord, 72) + dstep_read_cost(t, st, ord, 88);
      if (c < best_c) {
        best_c = c;
        for (int j = 0; j < 4; ++j) best[j] = ord[j];
      }
    }
    int y[4] = {}, x[4] = {}, lv[4] = {};
    for (int j = 0; j < 4; ++j) {
      y[j] = pair ? t.pty[st][best[j]] : t.sty[st - t.np][best[j]];
      x[j] = pair ? t.ptx[st][best[j]] : t.stx[st - t.np][best[j]];
      lv[j] = pair ? t.plive[st][best[j]] : t.slive[st - t.np][best[j]];
    }
    for (int j = 0; j < 4; ++j) {
      if (pair) {
        t.pty[st][j] = y[j];
        t.ptx[st][j] = x[j];
        t.plive[st][j] = lv[j];
      } else {
        t.sty[st - t.np][j] = y[j];
        t.stx[st - t.np][j] = x[j];
        t.slive[st - t.np][j] = lv[j];
      }
    }
  }
  return t;
}

// 2-D tables: the skewed map, taps ordered for it.
constexpr DTable make_dtable(int R, int shape) {
  DTable t = make_dtable_spans(spans_2d(R, shape), -1);
  t.map = 1;
  return order_dtable(t);
}

// Number of k-steps of a dilated table.
constexpr int dtable_ksteps(const DTable &t) { return 2 * t.np + t.ns; }

// Address plan of a dilated table's reads at row pitch px: read i (pairs first, then
// singles) is at lane-invariant distance delta[i] bytes (>= 0) from read rep[i], so one
// vector address per representative serves all reads of its class through the LDS
// instruction's immediate offset.  Two reads share a class when their per-lane cell
// offsets differ by the same amount in all four k lane groups.
struct RdPlan {
  int n, nrep;
  int rep[40], delta[40];
};

constexpr int dtable_cell(const DTable &t, int i, int q, int px) {
  return i < t.np ? t.pty[i][q] * px + t.ptx[i][q] : t.sty[i - t.np][q] * px + t.stx[i - t.np][q];
}

constexpr RdPlan make_rdplan(const DTable &t, int px) {
  RdPlan p{};
  p.n = t.np + t.ns;
  for (int i = 0; i < p.n; ++i) {
    p.rep[i] = i;
    for (int k = 0; k < i && p.rep[i] == i; ++k) {
      if (p.rep[k] != k) continue;
      const int d = dtable_cell(t, i, 0, px) - dtable_cell(t, k, 0, px);
      bool same = true;
      for (int q = 1; q < 4; ++q) same = same && dtable_cell(t, i, q, px) - dtable_cell(t, k, q, px) == d;
      if (same) p.rep[i] = k;
    }
  }
  // re-base every class on its member with the lowest address
  for (int i = 0; i < p.n; ++i) {
    if (p.rep[i] != i) continue;
    int lo = i;
    for (int k = i; k < p.n; ++k)
      if (p.rep[k] == i && dtable_cell(t, k, 0, px) < dtable_cell(t, lo, 0, px)) lo = k;
    for (int k = i; k < p.n; ++k)
      if (p.rep[k] == i) p.rep[k] = -1 - lo;
  }
  for (int i = 0; i < p.n; ++i) {
    p.rep[i] = -1 - p.rep[i];
    p.delta[i] = (dtable_cell(t, i, 0, px) - dtable_cell(t, p.rep[i], 0, px)) * 8;
  }
  p.nrep = 0;
  for (int i = 0; i < p.n; ++i) p.nrep += p.rep[i] == i ? 1 : 0;
  return p;
}

}  // namespace cs

namespace cs {

// ---- chunk table of the split-precision 2-D kernels (v_mfma_f32_16x16x32, 16 x 16 blocks) ----
//
// D[m][n] = sum_k A[m][k] B[k][n] with the weights as A and the tile as B.  Output
// (m, n) of a block at (y, x) is the cell (y, x) + u_n + v_m with anchors
// u_n = (2 (n & 7), 8 (n >> 3)) and offsets v_m = (m >> 3, m & 7), a bijection onto the
// 16 x 16 block.  K runs over the dilated footprint T = {d + v : d a tap, v in V},
//   A[m][t] = W[t - v_m] (0 off the stencil),   B[t][n] = tile[(y, x) + u_n + t],
// grouped into chunks of 8 consecutive columns t = (ty, tx0 .. tx0 + 7): lane l holds
// B[k][n] for n = l & 15 and the 8 k of chunk l >> 4 of the instruction -- one 16-byte
// LDS read -- and D[m][n] for m = 4 (l >> 4) + 0..3, i.e. 4 adjacent columns of one row.
// Chunks start at columns tx0 = c0 (mod 8), c0 in {0, 4} (fewest chunks), so that both
// the 16-byte reads and the 8-byte writes of the 4 adjacent results are aligned when
// block columns satisfy x + CS + c0 = 0 (mod 8).  box2d1r: 8 chunks (2 MFMAs per slice
// product per 256 outputs), box2d3r: 16 (4 MFMAs), star2d3r: 12 (3 MFMAs).
struct CTable {
  int map;           // anchor map of the 16 x 16 blocks (split_arow / split_aoct): 0 plain, 1 skew
  int n;             // chunks, a multiple of 4 (4 per MFMA)
  int c0;            // chunk start residue mod 8
  int ty[64], tx[64], live[64];
  int tmin, tmax;    // rows read relative to a block row anchor 2 (n & 7)
  int cmin, cmax;    // first and last chunk start column
};

// Chunk table of the footprint `sp` with chunk starts c0 (mod 8).
constexpr CTable make_ctable_spans(const Spans &sp, int c0) {
  const int R = sp.R;
  CTable t{};
  t.c0 = c0;
  int n = 0;
  for (int ty = -R; ty <= R + 1; ++ty) {
    int lo = 1 << 20, hi = -(1 << 20);
    for (int vy = 0; vy <= 1; ++vy) {
      const int dy = ty - vy;
      if (dy < -R || dy > R) continue;
      const int a = sp.lo[dy + R], b = sp.hi[dy + R];
      if (a > b) continue;
      lo = a < lo ? a : lo;
      hi = b + 7 > hi ? b + 7 : hi;
    }
    if (lo > hi) continue;
    for (int a = lo - (((lo - c0) % 8) + 8) % 8; a <= hi; a += 8) {
      t.ty[n] = ty;
      t.tx[n] = a;
      t.live[n] = 1;
      ++n;
    }
  }
  const int n4 = (n + 3) / 4 * 4;
  for (int e = n; e < n4; ++e) {  // padding chunks repeat the first chunk with zero weight
    t.ty[e] = t.ty[0];
    t.tx[e] = t.tx[0];
    t.live[e] = 0;
  }
  t.n = n4;
  t.tmin = 1 << 20;
  t.tmax = -(1 << 20);
  t.cmin = 1 << 20;
  t.cmax = -(1 << 20);
  for (int e = 0; e < n4; ++e) {
    t.tmin = t.ty[e] < t.tmin ? t.ty[e] : t.tmin;
    t.tmax = t.ty[e] > t.tmax ? t.ty[e] : t.tmax;
    t.cmin = t.tx[e] < t.cmin ? t.tx[e] : t.cmin;
    t.cmax = t.tx[e] > t.cmax ? t.tx[e] : t.cmax;
  }
  return t;
}

// Row-major chunk table of a 2-D stencil (the c0 with fewer chunks); make_ctable below
// orders its chunks for the LDS reads.
constexpr CTable make_ctable_rows(int R, int shape) {
  const CTable a = make_ctable_spans(spans_2d(R, shape), 0), b = make_ctable_spans(spans_2d(R, shape), 4);
  return b.n < a.n ? b : a;
}

// Address plan of a split table's B reads at row pitch px (elements of esz bytes): read i
// (instruction i, chunk 4 i + g in k lane group g) is at lane-invariant distance
// delta[i] bytes (>= 0) from read rep[i], as RdPlan above for the FP64 tables.
constexpr int ctable_cell(const CTable &t, int i, int g, int px) { return t.ty[4 * i + g] * px + t.tx[4 * i + g]; }

constexpr RdPlan make_splan(const CTable &t, int px, int esz) {
  RdPlan p{};
  p.n = t.n / 4;
  for (int i = 0; i < p.n; ++i) {
    p.rep[i] = i;
    for (int k = 0; k < i && p.rep[i] == i; ++k) {
      if (p.rep[k] != k) continue;
      const int d = ctable_cell(t, i, 0, px) - ctable_cell(t, k, 0, px);
      bool same = true;
      for (int g = 1; g < 4; ++g) same = same && ctable_cell(t, i, g, px) - ctable_cell(t, k, g, px) == d;
      if (same) p.rep[i] = k;
    }
  }
  for (int i = 0; i < p.n; ++i) {
    if (p.rep[i] != i) continue;
    int lo = i;
    for (int k = i; k < p.n; ++k)
      if (p.rep[k] == i && ctable_cell(t, k, 0, px) < ctable_cell(t, lo, 0, px)) lo = k;
    for (int k = i; k < p.n; ++k)
      if (p.rep[k] == i) p.rep[k] = -1 - lo;
  }
  for (int i = 0; i < p.n; ++i) {
    p.rep[i] = -1 - p.rep[i];
    p.delta[i] = (ctable_cell(t, i, 0, px) - ctable_cell(t, p.rep[i], 0, px)) * esz;
  }
  p.nrep = 0;
  for (int i = 0; i < p.n; ++i) p.nrep += p.rep[i] == i ? 1 : 0;
  return p;
}

}  // namespace cs

namespace cs {

// ---- dilated 4x4x4 tables of the 3-D kernels: one 2-D table per plane dz ----
// Every plane's table uses the same pair parity (a block's column alignment holds for all
// dz); the parity with the fewest k-steps (then LDS reads) over all planes is chosen.
struct DTable3 {
  int R, pp;
  DTable t[7];  // plane dz at t[dz + R]
};

constexpr Spans spans_3d(int R, int shape, int dz) {
  Spans s{};
  s.R = R;
  for (int dy = -R; dy <= R; ++dy) row_span_3d(R, shape, dz, dy, s.lo[dy + R], s.hi[dy + R]);
  return s;
}

constexpr DTable3 make_dtable3(int R, int shape) {
  DTable3 best{};
  int best_ks = 1 << 30, best_rd = 1 << 30;
  for (int pp = 0; pp < 2; ++pp) {
    DTable3 c{};
    c.R = R;
    c.pp = pp;
    int ks = 0, rd = 0;
    for (int dz = -R; dz <= R; ++dz) {
      c.t[dz + R] = make_dtable_spans(spans_3d(R, shape, dz), pp);
      ks += dtable_ksteps(c.t[dz + R]);
      rd += c.t[dz + R].np + c.t[dz + R].ns;
    }
    if (ks < best_ks || (ks == best_ks && rd < best_rd)) {
      best = c;
      best_ks = ks;
      best_rd = rd;
    }
  }
  return best;
}

constexpr int dtable3_ksteps(const DTable3 &t) {
  int n = 0;
  for (int dz = -t.R; dz <= t.R; ++dz) n += dtable_ksteps(t.t[dz + t.R]);
  return n;
}

}  // namespace cs

namespace cs {

// ---- LDS bank-conflict model of gfx950 (measured) ----
//
// Fitted to the measured cost of 330 lane-address patterns (strides, grouped rows, the
// kernels' own tables at every pitch):
//   * 16-byte reads: 64 banks of 4 bytes, four passes of 16 lanes; lane l is served in
//     pass (l >> 5, parity of bits 2..4 of l);
//   * 8-byte reads: 64 banks, two passes of 32 consecutive lanes;
//   * 8-byte writes: 32 banks, four passes of 16 consecutive lanes;
// a pass takes as many cycles as the largest number of dwords that fall in one bank.
// The model's cycle counts match the measured ones to a few percent (plus a fixed
// overhead of about 0.4 cycle per pass).  Addresses are bytes relative to any base.
constexpr int lds_cost(const int (&addr)[64], int bytes, bool write) {
  const int banks = write ? 32 : 64;
  int cnt[4][64] = {};
  for (int l = 0; l < 64; ++l) {
    int pass = 0;
    if (write)
      pass = l >> 4;
    else if (bytes == 16)
      pass = ((l >> 5) << 1) | (((l >> 2) ^ (l >> 3) ^ (l >> 4)) & 1);
    else
      pass = l >> 5;
    for (int k = 0; k < bytes / 4; ++k) {
      const int d = (addr[l] + 4 * k) / 4;
      ++cnt[pass][((d % banks) + banks) % banks];
    }
  }
  int cost = 0;
  for (int p = 0; p < 4; ++p) {
    int worst = 0;
    for (int b = 0; b < banks; ++b) worst = cnt[p][b] > worst ? cnt[p][b] : worst;
    cost += worst;
  }
  return cost;
}

// Workgroups per CU that an LDS footprint of `bytes` allows (160 KB per CU).
constexpr int lds_wgs(long bytes) { return bytes > 0 ? (int)(163840 / bytes) : 64; }

// LDS cost of one 8 x 8 block of a dilated table (its pair and single reads) plus the
// write-back of its results, at row pitch px doubles.
constexpr int dtable_lds_cost(const DTable &t, int px) {
  int cost = 0;
  int a[64] = {};
  const int ident[4] = {0, 1, 2, 3};
  for (int ps = 0; ps < t.np + t.ns; ++ps) cost += dstep_read_cost(t, ps, ident, px);
  for (int l = 0; l < 64; ++l) {  // D_b[i][n], i = l >> 4: cell (2 row + (n >> 1), 2 col + (n & 1))
    const int b = (l >> 2) & 3, n = l & 3, i = l >> 4;
    a[l] = ((2 * dil_row(t.map, b, i) + (n >> 1)) * px + 2 * dil_col(t.map, b, i) + (n & 1)) * 8;
  }
  return cost + lds_cost(a, 8, true);
}

// Pitch (doubles, >= px0, even, < px0 + 32) with the lowest cost over the given tables
// among the pitches that keep the most workgroups per CU; the kernel's LDS footprint is
// bpx * pitch + fixed bytes.
constexpr int best_pitch_dtables(const DTable *ts, int nt, int px0, long bpx, long fixed) {
  const int p0 = px0 + (px0 & 1);
  const int wg0 = lds_wgs(bpx * p0 + fixed);
  int best = p0, best_cost = 1 << 30;
  for (int px = p0; px < p0 + 32; px += 2) {
    if (lds_wgs(bpx * px + fixed) < wg0) break;
    int c = 0;
    for (int i = 0; i < nt; ++i) c += dtable_lds_cost(ts[i], px);
    if (c < best_cost) {
      best_cost = c;
      best = px;
    }
  }
  return best;
}

// Anchors of the split form's 16 x 16 blocks.  D column n of an MFMA (lane l: n = l & 15)
// is the 2 x 8 cell group at row pair split_arow(n) and column octet split_aoct(n) of the
// block; D row m = 4 (l >> 4) + v is the cell (m >> 3, m & 7) inside it.  Any bijection of
// n onto the 8 x 2 groups computes the same stencil (the A fragments only depend on m).
// Map 0 is n & 7, n >> 3 (the 3-D kernels); map 1 flips the octet of the row pairs 4..7,
// which with the chunk order of order_ctable makes the 16-byte B reads conflict-free and
// halves the conflicts of the 8-byte result writes at pitches of 8 (mod 16) elements.
constexpr int split_arow(int map, int n) { return (void)map, n & 7; }
constexpr int split_aoct(int map, int n) { return map == 1 ? ((n >> 3) ^ ((n >> 2) & 1)) : (n >> 3); }

// LDS cost of one 16 x 16 block of the split form at pitch px (elements of `esz` bytes):
// one 16-byte read per slice and instruction, one 8-byte write per slice.
constexpr int ctable_lds_cost(const CTable &t, int px, int esz) {
  int cost = 0;
  int a[64] = {};
  for (int i = 0; i < t.n / 4; ++i) {
    for (int l = 0; l < 64; ++l) {
      const int g = l >> 4, n = l & 15, c = 4 * i + g;
      a[l] = ((2 * split_arow(t.map, n) + t.ty[c] + 16) * px + 8 * split_aoct(t.map, n) + t.tx[c] + 16) * esz;
    }
    cost += lds_cost(a, 16, false);
  }
  for (int l = 0; l < 64; ++l) {
    const int g = l >> 4, n = l & 15;
    a[l] = ((2 * split_arow(t.map, n) + (g >> 1)) * px + 8 * split_aoct(t.map, n) + 4 * (g & 1)) * esz;
  }
  return cost + lds_cost(a, 8, true);
}

// Pitch (elements, a multiple of 8 >= px0) with the lowest cost among those that keep
// the most workgroups per CU (LDS footprint bpx * pitch + fixed bytes).
constexpr int best_pitch_ctable(const CTable &t, int px0, int esz, long bpx, long fixed) {
  const int p0 = (px0 + 7) / 8 * 8;
  const int wg0 = lds_wgs(bpx * p0 + fixed);
  int best = p0, best_cost = 1 << 30;
  for (int px = p0; px < p0 + 64; px += 8) {
    if (lds_wgs(bpx * px + fixed) < wg0) break;
    const int c = ctable_lds_cost(t, px, esz);
    if (c < best_cost) {
      best_cost = c;
      best = px;
    }
  }
  return best;
}

// ---- chunk order of the split tables ----
//
// The lanes of an MFMA's four k groups read chunks 4 i .. 4 i + 3; a 16-byte read is served
// in four passes of 16 lanes (lds_cost), and two lanes of one pass whose chunks lie in the
// same row 8 columns apart read the same 16 bytes from the same banks.  Row-major order puts
// such chunks in k groups 0 and 1 (the same passes); ordered as (ty, a), (ty + 1, a),
// (ty, a + 8), (ty + 1, a + 8) they fall in different passes, and at pitches of 16 (mod 32)
// elements every read of box2d3r, star2d3r and box2d1r is conflict-free in the model
// (4 cycles instead of 8).  order_ctable finds such orders: instruction by instruction,
// the first unused chunk takes k group 0 and every choice of three unused chunks for groups
// 1..3 is tried at pitch px, the cheapest kept (stopping at a conflict-free choice).
constexpr int kOrderPitch = 112;  // 2-byte slices: 16 (mod 32) elements

constexpr int cinstr_cost(const int (&ty)[4], const int (&tx)[4], int px, int esz, int map) {
  int a[64] = {};
  for (int l = 0; l < 64; ++l) {
    const int g = l >> 4, n = l & 15;
    a[l] = ((2 * split_arow(map, n) + ty[g] + 16) * px + 8 * split_aoct(map, n) + tx[g] + 16) * esz;
  }
  return lds_cost(a, 16, false);
}

constexpr CTable order_ctable(const CTable &t, int px, int esz) {
  CTable o = t;
  bool used[64] = {};
  for (int i = 0; i < t.n / 4; ++i) {
    int c0 = 0;
    while (used[c0]) ++c0;
    used[c0] = true;
    int best = 1 << 30, b1 = -1, b2 = -1, b3 = -1;
    for (int x1 = 0; x1 < t.n && best > 4; ++x1) {
      if (used[x1]) continue;
      for (int x2 = 0; x2 < t.n && best > 4; ++x2) {
        if (used[x2] || x2 == x1) continue;
        for (int x3 = 0; x3 < t.n && best > 4; ++x3) {
          if (used[x3] || x3 == x1 || x3 == x2) continue;
          const int ty[4] = {t.ty[c0], t.ty[x1], t.ty[x2], t.ty[x3]};
          const int tx[4] = {t.tx[c0], t.tx[x1], t.tx[x2], t.tx[x3]};
          const int c = cinstr_cost(ty, tx, px, esz, t.map);
          if (c < best) {
            best = c;
            b1 = x1;
            b2 = x2;
            b3 = x3;
          }
        }
      }
    }
    used[b1] = used[b2] = used[b3] = true;
    const int sel[4] = {c0, b1, b2, b3};
    for (int g = 0; g < 4; ++g) {
      o.ty[4 * i + g] = t.ty[sel[g]];
      o.tx[4 * i + g] = t.tx[sel[g]];
      o.live[4 * i + g] = t.live[sel[g]];
    }
  }
  return o;
}

// Chunk table of the split-precision 2-D kernels (and their host-side A fragments): the
// skewed anchor map, chunks ordered for it at a pitch of 8 (mod 16) elements.
constexpr int kOrderPitchSkew = 88;
constexpr CTable make_ctable(int R, int shape) {
  CTable t = make_ctable_rows(R, shape);
  t.map = 1;
  return order_ctable(t, kOrderPitchSkew, 2);
}

// One evaluation per stencil class for the kernels' compile-time tables.
template <int R, int SH>
struct CTab2 {
  static constexpr CTable T = make_ctable(R, SH);
};

// ---- chunk tables of the split-precision 3-D kernels: one CTable per plane dz ----
// Every plane's chunks start at the same residue c0 (a block's column alignment then
// holds for each dz); the c0 with the fewest chunks over all planes is chosen.  Each
// plane's chunks are padded to whole MFMAs on their own: an MFMA reads one plane.
struct CTable3 {
  int R, c0;
  CTable t[7];  // plane dz at t[dz + R]
};

constexpr CTable3 make_ctable3(int R, int shape) {
  CTable3 best{};
  int best_n = 1 << 30;
  for (int c0 = 0; c0 <= 4; c0 += 4) {
    CTable3 c{};
    c.R = R;
    c.c0 = c0;
    int n = 0;
    for (int dz = -R; dz <= R; ++dz) {
      c.t[dz + R] = make_ctable_spans(spans_3d(R, shape, dz), c0);
      n += c.t[dz + R].n;
    }
    if (n < best_n) {
      best = c;
      best_n = n;
    }
  }
  for (int dz = -R; dz <= R; ++dz) best.t[dz + R] = order_ctable(best.t[dz + R], kOrderPitch, 2);
  return best;
}

template <int R, int SH>
struct CTab3 {
  static constexpr CTable3 T = make_ctable3(R, SH);
};

// As best_pitch_ctable, with the reads of every plane counted.
constexpr int best_pitch_ctable3(const CTable3 &t, int px0, int esz, long bpx, long fixed) {
  const int p0 = (px0 + 7) / 8 * 8;
  const int wg0 = lds_wgs(bpx * p0 + fixed);
  int best = p0, best_cost = 1 << 30;
  for (int px = p0; px < p0 + 64; px += 8) {
    if (lds_wgs(bpx * px + fixed) < wg0) break;
    int c = 0;
    for (int dz = -t.R; dz <= t.R; ++dz) c += ctable_lds_cost(t.t[dz + t.R], px, esz);
    if (c < best_cost) {
      best_cost = c;
      best = px;
    }
  }
  return best;
}

}  // namespace cs
