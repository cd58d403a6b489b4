// k3d_reg.hip -- register-streaming FP64 3-D single-step kernels for gfx950 (radius 1, the
// 27-point box and the 7-point star).
//
// Replaces gpu_box_3d1r_kernel and gpu_star_3d1r_step3_kernel (reference
// src/3d/gpu_box.cu:186-238, src/3d/gpu_star.cu) for physical single steps.  A wave owns a
// tile of 64 PX columns by BY rows and streams its z range, one input plane per iteration,
// with no LDS and no barrier: lane l holds columns l PX .. l PX + PX - 1 of the BY + 2 input
// rows a plane contributes (the tile's rows and one on either side), the column
// neighbours beyond a lane's run come from lanes l -+ 1 (v_mov_b32_dpp wave shifts), and
// the columns just outside the tile from one extra load per row that lanes 0 and 63 hold
// (the shift's "old" operand, taken where the shifted lane does not exist).  An incoming
// plane z adds its dz = +1 taps to output plane z - 1 (then complete and stored), its dz = 0
// taps to plane z and starts plane z + 1 with its dz = -1 taps: three partial planes per
// wave rotate every three iterations.  One step per launch, so cells outside the interior
// are simply not stored.  Sum order per output: dz, dy, dx ascending (the reference's).
#include "cs_internal.h"

namespace cs {

// lane l receives lane l - 1's value, lane 0 keeps `old`
__device__ __forceinline__ double lane_from_lower_or(double v, double old) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), 0x138, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// lane l receives lane l + 1's value, lane 63 keeps `old`
__device__ __forceinline__ double lane_from_upper_or(double v, double old) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), 0x130, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), 0x130, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int PX, int BY>
struct Reg3D {
  static constexpr int OX = 64 * PX;  // output columns per wave
  static constexpr int NR = BY + 2;   // input rows per plane
};

template <int SH, int PX, int BY, int RING, bool EDGE>
struct Sweep3D {
  using C = Reg3D<PX, BY>;
  static constexpr int NR = C::NR;
  static constexpr bool BOX = SH == SHAPE3_BOX;
  static constexpr int PER = RING % 3 == 0 ? RING : 3 * RING;

  const double *__restrict__ in;
  double *__restrict__ out;
  double w[27];
  int64_t nz, ny, nx, ez, ey, ex, oz, oy, ox;
  int64_t y0, cx, xe;  // first output row; this lane's first column; its extra column
  int lane;
  int64_t colc[PX], xec;  // checked path: the columns clamped into the array
  bool colout[PX];
  double ring[RING][NR][PX], ext[RING][NR];
  double acc[3][BY][PX];

  // input plane z, rows y0 - 1 .. y0 + BY: the lane's run and its extra cell (lane 0: the
  // column left of the tile, lane 63: the column right of it).  The checked path clamps
  // plane, row and column into the array: a cell outside it is at least one cell from the
  // interior (the halo is at least one cell wide), so whatever it holds reaches only outputs
  // that are never stored -- and no select waits on the load.
  __device__ __forceinline__ void load_plane(int64_t z, double (&q)[NR][PX], double (&e)[NR]) {
    const int64_t az = oz + z;
    if constexpr (!EDGE) {
      const double *p = in + ((az * ey) + oy + y0 - 1) * ex + ox;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
#pragma unroll
        for (int j = 0; j < PX; ++j) q[r][j] = p[r * ex + cx + j];
        e[r] = p[r * ex + xe];
      }
    } else {
      const int64_t cz = az < 0 ? 0 : az >= ez ? ez - 1 : az;  // wave-uniform
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int64_t ay = oy + y0 - 1 + r;
        const int64_t cy = ay < 0 ? 0 : ay >= ey ? ey - 1 : ay;
        const double *p = in + (cz * ey + cy) * ex;
#pragma unroll
        for (int j = 0; j < PX; ++j) q[r][j] = p[colc[j]];
        e[r] = p[xec];
      }
    }
  }

  template <int K>
  __device__ __forceinline__ void phase(int64_t it, int64_t zs, int64_t z0, int64_t z1) {
    const int64_t z = zs + it + K;
    load_plane(z + RING - 1, ring[(K + RING - 1) % RING], ext[(K + RING - 1) % RING]);
    const double(&q)[NR][PX] = ring[K % RING];
    const double(&e)[NR] = ext[K % RING];
    // the rows extended by one cell on either side: qe[r][1 + c], c = -1 .. PX
    double qe[NR][PX + 2];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
#pragma unroll
      for (int j = 0; j < PX; ++j) qe[r][1 + j] = q[r][j];
      qe[r][0] = lane_from_lower_or(q[r][PX - 1], e[r]);
      qe[r][PX + 1] = lane_from_upper_or(q[r][0], e[r]);
    }
    // output plane z - dz sits in slot (z - dz - zs) mod 3 = (K - dz) mod 3
#pragma unroll
    for (int dz = 1; dz >= -1; --dz) {
      double(&T)[BY][PX] = acc[((K - dz) % 3 + 3) % 3];
#pragma unroll
      for (int i = 0; i < BY; ++i)
#pragma unroll
        for (int j = 0; j < PX; ++j) {
          double t = dz == -1 ? 0.0 : T[i][j];
          bool first = dz == -1;
#pragma unroll
          for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
              if (!BOX && (dz != 0) + (dy != 0) + (dx != 0) > 1) continue;
              const double wv = w[((dz + 1) * 3 + dy + 1) * 3 + dx + 1], v = qe[i + 1 + dy][1 + j + dx];
              t = first ? wv * v : __builtin_fma(wv, v, t);
              first = false;
            }
          T[i][j] = t;
        }
      if (dz < 1) {
#pragma unroll
        for (int i = 0; i < BY; ++i)
#pragma unroll
          for (int j = 0; j < PX; ++j) asm volatile("" : "+v"(T[i][j]));
      }
    }
    // plane z - 1 is complete
    const int64_t zo = z - 1;
    store_plane(zo, zo >= z0 && zo < z1, acc[(K - 1 + 3) % 3]);
  }

  // Output plane zo's rows by buffer stores over each row of the tile: a plane outside
  // [z0, z1) or a row outside the interior gets an empty range and a cell outside it an
  // offset past the range, so every lane issues the same stores on every plane (no branch
  // in the loop, and the compiler's waits count the stores exactly).
  __device__ __forceinline__ void store_plane(int64_t zo, bool ok, const double (&A)[BY][PX]) {
    const int base = lane * PX * 8;
#pragma unroll
    for (int i = 0; i < BY; ++i) {
      const bool rok = ok && (!EDGE || y0 + i < ny);
      const __amdgpu_buffer_rsrc_t rs =
          buffer_over(out + ((oz + zo) * ey + oy + y0 + i) * ex + ox + cx - lane * PX, rok ? C::OX * 8 : 0);
      if constexpr (!EDGE && PX % 2 == 0) {
#pragma unroll
        for (int j = 0; j < PX; j += 2) buffer_store2(rs, base + j * 8, A[i][j], A[i][j + 1]);
      } else {
#pragma unroll
        for (int j = 0; j < PX; ++j) buffer_store1(rs, !EDGE || colout[j] ? base + j * 8 : kBufferDrop, A[i][j]);
      }
    }
  }

  template <int K>
  __device__ __forceinline__ void phases(int64_t it, int64_t zs, int64_t z0, int64_t z1) {
    if constexpr (K < PER) {
      phase<K>(it, zs, z0, z1);
      phases<K + 1>(it, zs, z0, z1);
    }
  }

  __device__ __forceinline__ void run(int64_t z0, int64_t z1) {
    const int64_t zs = z0 - 1;
    const int64_t iters = (z1 - z0) + 2;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < BY; ++i)
#pragma unroll
        for (int j = 0; j < PX; ++j) acc[k][i][j] = 0.0;
    // The loop head's wait is one instruction for the first entry and every later one, so
    // the prologue issues the loop's own pattern: each plane's loads, in order (no
    // reordering across the scheduling barrier), then one plane of dropped stores.
#pragma unroll
    for (int k = 0; k < RING - 1; ++k) {
      load_plane(zs + k, ring[k], ext[k]);
      const double none[BY][PX] = {};
      store_plane(zs, false, none);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int64_t it = 0; it < iters; it += PER) phases<0>(it, zs, z0, z1);
  }
};

template <int SH, int PX, int BY, int RING, bool EDGE>
__device__ __forceinline__ void sweep3(const double *__restrict__ in, double *__restrict__ out,
                                       const double *__restrict__ wts, const Geo &g, int64_t x0, int64_t y0,
                                       int64_t z0, int64_t z1, int lane) {
  Sweep3D<SH, PX, BY, RING, EDGE> S;
  S.in = in;
  S.out = out;
#pragma unroll
  for (int i = 0; i < 27; ++i) S.w[i] = wts[i];
  S.nz = g.n[0];
  S.ny = g.n[1];
  S.nx = g.n[2];
  S.ez = g.ext[0];
  S.ey = g.ext[1];
  S.ex = g.ext[2];
  S.oz = g.org[0];
  S.oy = g.org[1];
  S.ox = g.org[2];
  S.y0 = y0;
  S.lane = lane;
  S.cx = x0 + (int64_t)lane * PX;
  // lane 0 holds the column left of the tile, lane 63 the one right of it, the others
  // their own first column (never used)
  S.xe = lane == 0 ? x0 - 1 : lane == 63 ? x0 + 64 * PX : S.cx;
  S.xec = S.ox + S.xe < 0 ? 0 : S.ox + S.xe >= S.ex ? S.ex - 1 : S.ox + S.xe;
#pragma unroll
  for (int j = 0; j < PX; ++j) {
    const int64_t c = S.cx + j;
    S.colc[j] = S.ox + c < 0 ? 0 : S.ox + c >= S.ex ? S.ex - 1 : S.ox + c;
    S.colout[j] = c >= 0 && c < S.nx;
  }
  S.run(z0, z1);
}

template <int SH, int PX, int BY, int RING, int NW, int WPE>
__global__ __launch_bounds__(NW * 64, WPE) void k3d_reg(const double *__restrict__ in, double *__restrict__ out,
                                                        const double *__restrict__ wts, Geo g, Reg3Split sp) {
  using C = Reg3D<PX, BY>;
  constexpr int PER = RING % 3 == 0 ? RING : 3 * RING;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * NW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t yt = wid % sp.ytiles, rest = wid / sp.ytiles;
  const int64_t strip = rest % sp.strips, zch = rest / sp.strips;
  if (zch >= sp.zchunks) return;
  const int64_t z0 = zch * sp.zc;
  if (z0 >= g.n[0]) return;
  const int64_t z1 = z0 + sp.zc < g.n[0] ? z0 + sp.zc : g.n[0];
  const int64_t y0 = yt * BY, x0 = strip * C::OX;
  // the plain path: every cell the sweep loads (planes z0 - 1 .. z1 + period + RING - 2,
  // rows y0 - 1 .. y0 + BY, columns x0 - 1 .. x0 + 64 PX) and stores is interior
  if (x0 >= 1 && x0 + C::OX + 1 <= g.n[2] && y0 >= 1 && y0 + BY + 1 <= g.n[1] && z0 >= 1 &&
      z1 + PER + RING <= g.n[0])
    sweep3<SH, PX, BY, RING, false>(in, out, wts, g, x0, y0, z0, z1, lane);
  else
    sweep3<SH, PX, BY, RING, true>(in, out, wts, g, x0, y0, z0, z1, lane);
}

// ---- KF fused steps (KF = 2, 3; zero halo: the plan fuses only when both buffers carry a
// zero halo)
//
// A wave owns a strip of 64 PX columns whose outer KF columns on either side are recomputed
// by the neighbouring strips (64 PX - 2 KF outputs), and BY output rows.  Input planes
// stream in z: rows y0 - KF .. y0 + BY + KF - 1 of plane z add their taps to three partial
// planes of level 1 (rows y0 - KF + 1 .. y0 + BY + KF - 2, as the single-step sweep does);
// the level-1 plane z - 1 then complete is masked to zero outside the interior (the halo
// of the intermediate step) and adds its taps to three partial planes of level 2 (two rows
// fewer), and so on: level l's plane z - l is complete after plane z arrives, and level
// KF's is stored.  Column neighbours come from the lanes l -+ 1; lane 0 and lane 63 read
// zero there, which only spoils the outer columns' values.
template <int PX, int BY, int KF>
struct Fuse3D {
  static constexpr int OX = 64 * PX - 2 * KF;  // output columns per wave
  static constexpr int NR = BY + 2 * KF;       // input rows per plane
  // rows of level l (1 .. KF): BY + 2 (KF - l)
  static constexpr int rows(int l) { return BY + 2 * (KF - l); }
};

template <int SH, int PX, int BY, int KF, int RING, bool EDGE>
struct Sweep3F {
  using C = Fuse3D<PX, BY, KF>;
  static constexpr int NR = C::NR, NA = C::rows(1);
  static constexpr bool BOX = SH == SHAPE3_BOX;
  static constexpr int PER = RING % 3 == 0 ? RING : 3 * RING;

  const double *__restrict__ in;
  double *__restrict__ out;
  double w[27];
  int64_t nz, ny, nx, ez, ey, ex, oz, oy, ox;
  int64_t y0, xs, cx;  // first output row; the strip's first column; this lane's first column
  int lane;
  int dcol[PX];  // checked path: column j + dcol[j] of the lane's run is column j clamped
                 // into the array
  bool colin[PX], colout[PX];
  double ring[RING][NR][PX];
  // partial planes of level l + 1: a[l][slot][row], rows 0 .. rows(l + 1) - 1 in use
  double a[KF][3][NA][PX];

  // input plane z, rows y0 - KF .. y0 + BY + KF - 1 (the checked path clamps plane, row and
  // column into the array: a cell outside it is at least two cells from the interior, so it
  // reaches only level-1 cells outside the interior, which are masked)
  __device__ __forceinline__ void load_plane(int64_t z, double (&q)[NR][PX]) {
    const int64_t az = oz + z;
    if constexpr (!EDGE) {
      const double *p = in + ((az * ey) + oy + y0 - KF) * ex + ox + cx;
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int j = 0; j < PX; ++j) q[r][j] = p[r * ex + j];
    } else {
      const int64_t cz = az < 0 ? 0 : az >= ez ? ez - 1 : az;  // wave-uniform
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int64_t ay = oy + y0 - KF + r;
        const int64_t cy = ay < 0 ? 0 : ay >= ey ? ey - 1 : ay;
        const double *p = in + (cz * ey + cy) * ex + ox + cx;
#pragma unroll
        for (int j = 0; j < PX; ++j) q[r][j] = p[j + dcol[j]];
      }
    }
  }

  // the taps of one plane offset dz from rows src (extended: src[r][1 + c], c = -1 .. PX) to
  // NO rows of the partial plane T (row i takes source rows i .. i + 2); start: the first
  // contribution to T
  template <int NO, int NI, int NT>
  __device__ __forceinline__ void taps(const double (&src)[NI][PX + 2], double (&T)[NT][PX], int dz, bool start) {
#pragma unroll
    for (int i = 0; i < NO; ++i)
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        double t = start ? 0.0 : T[i][j];
        bool first = start;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            if (!BOX && (dz != 0) + (dy != 0) + (dx != 0) > 1) continue;
            const double wv = w[((dz + 1) * 3 + dy + 1) * 3 + dx + 1], v = src[i + 1 + dy][1 + j + dx];
            t = first ? wv * v : __builtin_fma(wv, v, t);
            first = false;
          }
        T[i][j] = t;
      }
  }

  template <int NO, int NT>
  __device__ __forceinline__ void pin(double (&T)[NT][PX]) {
#pragma unroll
    for (int i = 0; i < NO; ++i)
#pragma unroll
      for (int j = 0; j < PX; ++j) asm volatile("" : "+v"(T[i][j]));
  }

  template <int NI, int NS>
  __device__ __forceinline__ void extend(const double (&q)[NS][PX], double (&qe)[NI][PX + 2]) {
#pragma unroll
    for (int r = 0; r < NI; ++r) {
#pragma unroll
      for (int j = 0; j < PX; ++j) qe[r][1 + j] = q[r][j];
      qe[r][0] = lane_from_lower(q[r][PX - 1]);
      qe[r][PX + 1] = lane_from_upper(q[r][0]);
    }
  }

  // Level L (1 .. KF) takes its input plane p = z - (L - 1) (rows NI) and adds its taps to
  // its planes p + 1 (start), p and p - 1 (then complete); level L's plane q sits in slot
  // (q - zs) mod 3 = (K - (L - 1) - dz) mod 3 for q = p - dz.
  template <int K, int L, int NI, int NS>
  __device__ __forceinline__ void level(int64_t z, const double (&src)[NS][PX]) {
    constexpr int NO = C::rows(L);
    {
      double qe[NI][PX + 2];
      extend<NI, NS>(src, qe);
#pragma unroll
      for (int dz = 1; dz >= -1; --dz) {
        double(&T)[NA][PX] = a[L - 1][((K - (L - 1) - dz) % 3 + 3) % 3];
        taps<NO, NI, NA>(qe, T, dz, dz == -1);
        if (dz < 1) pin<NO, NA>(T);
      }
    }
    // level L's plane z - L is complete
    double(&P)[NA][PX] = a[L - 1][((K - L) % 3 + 3) % 3];
    if constexpr (L < KF) {
      // outside the interior it is the halo (zero)
      if constexpr (EDGE) {
        const int64_t p = z - L;
        const bool zin = p >= 0 && p < nz;  // wave-uniform
#pragma unroll
        for (int i = 0; i < NO; ++i) {
          const int64_t y = y0 - (KF - L) + i;
          const bool rin = zin && y >= 0 && y < ny;  // wave-uniform
#pragma unroll
          for (int j = 0; j < PX; ++j) P[i][j] = rin && colin[j] ? P[i][j] : 0.0;
        }
      }
      level<K, L + 1, NO, NA>(z, P);
    }
  }

  template <int K>
  __device__ __forceinline__ void phase(int64_t it, int64_t zs, int64_t z0, int64_t z1) {
    const int64_t z = zs + it + K;
    load_plane(z + RING - 1, ring[(K + RING - 1) % RING]);
    level<K, 1, NR, NR>(z, ring[K % RING]);
    // level KF's plane z - KF is complete
    const int64_t zo = z - KF;
    store_plane(zo, zo >= z0 && zo < z1, a[KF - 1][((K - KF) % 3 + 3) % 3]);
  }

  // Output plane zo's rows by buffer stores over each row of the strip: a plane outside
  // [z0, z1) or a row outside the interior gets an empty range and a cell outside the
  // outputs an offset past the range (no branch in the loop).
  __device__ __forceinline__ void store_plane(int64_t zo, bool ok, const double (&A)[NA][PX]) {
    const int base = lane * PX * 8;
#pragma unroll
    for (int i = 0; i < BY; ++i) {
      const bool rok = ok && (!EDGE || y0 + i < ny);
      const __amdgpu_buffer_rsrc_t rs =
          buffer_over(out + ((oz + zo) * ey + oy + y0 + i) * ex + ox + xs, rok ? 64 * PX * 8 : 0);
      if constexpr (!EDGE && PX % 2 == 0 && KF % 2 == 0) {
        // the plain path: the outputs start and end at even lane-run offsets (KF and
        // 64 PX - KF), so both cells of a pair are outputs or neither (on the checked path
        // the grid's last column can split a pair)
#pragma unroll
        for (int j = 0; j < PX; j += 2)
          buffer_store2(rs, colout[j] ? base + j * 8 : kBufferDrop, A[i][j], A[i][j + 1]);
      } else {
#pragma unroll
        for (int j = 0; j < PX; ++j) buffer_store1(rs, colout[j] ? base + j * 8 : kBufferDrop, A[i][j]);
      }
    }
  }

  template <int K>
  __device__ __forceinline__ void phases(int64_t it, int64_t zs, int64_t z0, int64_t z1) {
    if constexpr (K < PER) {
      phase<K>(it, zs, z0, z1);
      phases<K + 1>(it, zs, z0, z1);
    }
  }

  __device__ __forceinline__ void run(int64_t z0, int64_t z1) {
    const int64_t zs = z0 - KF;
    const int64_t iters = (z1 - z0) + 2 * KF;
#pragma unroll
    for (int l = 0; l < KF; ++l)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int j = 0; j < PX; ++j) a[l][k][i][j] = 0.0;
#pragma unroll
    for (int k = 0; k < RING - 1; ++k) {
      load_plane(zs + k, ring[k]);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int64_t it = 0; it < iters; it += PER) phases<0>(it, zs, z0, z1);
  }
};

template <int SH, int PX, int BY, int KF, int RING, bool EDGE>
__device__ __forceinline__ void sweep3f(const double *__restrict__ in, double *__restrict__ out,
                                        const double *__restrict__ wts, const Geo &g, int64_t xs, int64_t y0,
                                        int64_t z0, int64_t z1, int lane) {
  Sweep3F<SH, PX, BY, KF, RING, EDGE> S;
  S.in = in;
  S.out = out;
#pragma unroll
  for (int i = 0; i < 27; ++i) S.w[i] = wts[i];
  S.nz = g.n[0];
  S.ny = g.n[1];
  S.nx = g.n[2];
  S.ez = g.ext[0];
  S.ey = g.ext[1];
  S.ex = g.ext[2];
  S.oz = g.org[0];
  S.oy = g.org[1];
  S.ox = g.org[2];
  S.y0 = y0;
  S.xs = xs;
  S.lane = lane;
  S.cx = xs + (int64_t)lane * PX;
#pragma unroll
  for (int j = 0; j < PX; ++j) {
    const int64_t c = S.cx + j, d = (int64_t)lane * PX + j;
    S.dcol[j] = (int)(S.ox + c < 0 ? -(S.ox + c) : S.ox + c >= S.ex ? S.ex - 1 - (S.ox + c) : 0);
    S.colin[j] = c >= 0 && c < S.nx;
    S.colout[j] = d >= KF && d < KF + Fuse3D<PX, BY, KF>::OX && c < S.nx;  // c >= 0 wherever d >= KF
  }
  S.run(z0, z1);
}

template <int SH, int PX, int BY, int KF, int RING, int NW, int WPE>
__global__ __launch_bounds__(NW * 64, WPE) void k3f_reg(const double *__restrict__ in, double *__restrict__ out,
                                                        const double *__restrict__ wts, Geo g, Reg3Split sp) {
  using C = Fuse3D<PX, BY, KF>;
  constexpr int PER = RING % 3 == 0 ? RING : 3 * RING;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * NW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t yt = wid % sp.ytiles, rest = wid / sp.ytiles;
  const int64_t strip = rest % sp.strips, zch = rest / sp.strips;
  if (zch >= sp.zchunks) return;
  const int64_t z0 = zch * sp.zc;
  if (z0 >= g.n[0]) return;
  const int64_t z1 = z0 + sp.zc < g.n[0] ? z0 + sp.zc : g.n[0];
  const int64_t y0 = yt * BY, xs = strip * C::OX - KF;
  // the plain path: every cell the sweep loads (planes z0 - KF .. z1 + KF + period + RING -
  // 3, rows y0 - KF .. y0 + BY + KF - 1, the strip's columns) and every intermediate cell is
  // interior
  if (xs >= 0 && xs + 64 * PX <= g.n[2] && y0 >= KF && y0 + BY + KF <= g.n[1] && z0 >= KF &&
      z1 + KF + PER + RING - 2 <= g.n[0])
    sweep3f<SH, PX, BY, KF, RING, false>(in, out, wts, g, xs, y0, z0, z1, lane);
  else
    sweep3f<SH, PX, BY, KF, RING, true>(in, out, wts, g, xs, y0, z0, z1, lane);
}

template <int SH, int PX, int BY, int RING, int NW, int WPE>
struct Reg3Slots {
  static int64_t value;
};
template <int SH, int PX, int BY, int RING, int NW, int WPE>
int64_t Reg3Slots<SH, PX, BY, RING, NW, WPE>::value = 0;

template <int SH, int PX, int BY, int RING, int NW, int WPE>
static hipError_t prepare_reg3d() {
  int nb = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, (const void *)k3d_reg<SH, PX, BY, RING, NW, WPE>, NW * 64, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  Reg3Slots<SH, PX, BY, RING, NW, WPE>::value = (int64_t)(nb > 0 ? nb : 1) * NW * (cus > 0 ? cus : 1);
  return hipSuccess;
}

template <int SH, int PX, int BY, int RING, int NW, int WPE>
static hipError_t launch_reg3d(const double *in, double *out, const double *wts, const Geo &g, hipStream_t s) {
  using C = Reg3D<PX, BY>;
  using S = Reg3Slots<SH, PX, BY, RING, NW, WPE>;
  const int64_t slots = S::value > 0 ? S::value : 2048;
  const Reg3Split sp = reg3_split(g.n[0], g.n[1], g.n[2], C::OX, BY, slots);
  const int64_t waves = sp.strips * sp.ytiles * sp.zchunks;
  hipLaunchKernelGGL((k3d_reg<SH, PX, BY, RING, NW, WPE>), dim3((unsigned)((waves + NW - 1) / NW)), dim3(NW * 64), 0,
                     s, in, out, wts, g, sp);
  return hipGetLastError();
}

template <int SH, int PX, int BY, int KF, int RING, int NW, int WPE>
struct Fuse3Slots {
  static int64_t value;
};
template <int SH, int PX, int BY, int KF, int RING, int NW, int WPE>
int64_t Fuse3Slots<SH, PX, BY, KF, RING, NW, WPE>::value = 0;

template <int SH, int PX, int BY, int KF, int RING, int NW, int WPE>
static hipError_t prepare_fuse3d() {
  int nb = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, (const void *)k3f_reg<SH, PX, BY, KF, RING, NW, WPE>, NW * 64, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  Fuse3Slots<SH, PX, BY, KF, RING, NW, WPE>::value = (int64_t)(nb > 0 ? nb : 1) * NW * (cus > 0 ? cus : 1);
  return hipSuccess;
}

template <int SH, int PX, int BY, int KF, int RING, int NW, int WPE>
static hipError_t launch_fuse3d(const double *in, double *out, const double *wts, const Geo &g, hipStream_t s) {
  using C = Fuse3D<PX, BY, KF>;
  using S = Fuse3Slots<SH, PX, BY, KF, RING, NW, WPE>;
  const int64_t slots = S::value > 0 ? S::value : 2048;
  const Reg3Split sp = reg3_split(g.n[0], g.n[1], g.n[2], C::OX, BY, slots);
  const int64_t waves = sp.strips * sp.ytiles * sp.zchunks;
  hipLaunchKernelGGL((k3f_reg<SH, PX, BY, KF, RING, NW, WPE>), dim3((unsigned)((waves + NW - 1) / NW)),
                     dim3(NW * 64), 0, s, in, out, wts, g, sp);
  return hipGetLastError();
}

// Variant::mf 9 marks this family: by = rows per wave, bx = columns per wave, px = columns
// per lane, st = planes in the load ring; the weights are the dense 3 x 3 x 3 block.
#define CS_VR3D(ID, SH, PX, BY, RING, WPE)                                                                   \
  Variant {                                                                                                  \
    ID, 9, 1, SH, 1, BY, Reg3D<PX, BY>::OX, 4, 0, PX, RING, &launch_reg3d<SH, PX, BY, RING, 4, WPE>,        \
        &prepare_reg3d<SH, PX, BY, RING, 4, WPE>,                                                           \
        "k3d_reg<r1," #SH ",p" #PX ",by" #BY ",ring" #RING ",wpe" #WPE ">"                                  \
  }

#define CS_VR3F(ID, SH, PX, BY, RING, WPE, NW, KF)                                                           \
  Variant {                                                                                                  \
    ID, 9, 1, SH, KF, BY, Fuse3D<PX, BY, KF>::OX, NW, 0, PX, RING,                                           \
        &launch_fuse3d<SH, PX, BY, KF, RING, NW, WPE>, &prepare_fuse3d<SH, PX, BY, KF, RING, NW, WPE>,      \
        "k3f_reg<r1," #SH ",p" #PX ",by" #BY ",kf" #KF ",ring" #RING ",wpe" #WPE ",w" #NW ">"               \
  }

// ids 900000 + (star) * 10000 + PX * 1000 + BY * 10 + RING; the fused ones 1000000 (two
// steps) or 3000000 (three) + (8-wave workgroups) * 800000 + (star) * 100000 + PX * 10000 +
// waves per SIMD * 1000 + BY * 10 + RING
// Two fused steps, measured at 512^3, T = 32 (GStencil/s; the first of each shape is the
// box's default): box p1
// by8 ring 3 449-451 (ring 2 447, by10 437, by9 427, by6 373, p2 by4 383-386, p2 by3 349, p2
// by2 276-294, 3 waves per SIMD at by5 / by6 400 / 347), with 8-wave workgroups 440
// against 428 on one box; star p1 by8 ring 3
// 448-472 (8-wave workgroups 445 against 448, by10 449-453, by12 435, by6 371, p2 by4
// 389-414, p2 by2 327-405).  Three fused steps: star p1 by7 ring 2 621, by8 584 (11 spilled
// registers; 8-wave workgroups 585), by6 546-563, by4 463, p2 by4 spills; box by6 433 (equal
// to two steps: its 27 taps make it FP64-issue-bound), by4 279.  Four fused steps: star by4
// 498, by5 458 (the recomputed rows outgrow the saved bytes).  Also slower: the star's three
// steps with 8-wave workgroups (543 against 620), the box with a ring of 4 planes (416-418
// against 432).  Taller tiles recompute fewer level-1 rows ((BY + 2) / BY) and
// load fewer halo rows ((BY + 4) / BY) per output row; one column per lane leaves the
// registers for them.
static const Variant kVariantsReg3D[] = {
    CS_VR3F(1812083, SHAPE3_BOX, 1, 8, 3, 2, 8, 2),   CS_VR3F(1012083, SHAPE3_BOX, 1, 8, 3, 2, 4, 2),
    CS_VR3F(1012082, SHAPE3_BOX, 1, 8, 2, 2, 4, 2),
    CS_VR3F(1022043, SHAPE3_BOX, 2, 4, 3, 2, 4, 2),   CS_VR3F(1022023, SHAPE3_BOX, 2, 2, 3, 2, 4, 2),
    CS_VR3F(1112083, SHAPE3_STAR, 1, 8, 3, 2, 4, 2),  CS_VR3F(1112102, SHAPE3_STAR, 1, 10, 2, 2, 4, 2),
    CS_VR3F(1122043, SHAPE3_STAR, 2, 4, 3, 2, 4, 2),  CS_VR3F(1122022, SHAPE3_STAR, 2, 2, 2, 2, 4, 2),
    CS_VR3F(1912083, SHAPE3_STAR, 1, 8, 3, 2, 8, 2),
    // three fused steps: the star's default (box: equal to two steps)
    CS_VR3F(3112072, SHAPE3_STAR, 1, 7, 2, 2, 4, 3),  CS_VR3F(3112082, SHAPE3_STAR, 1, 8, 2, 2, 4, 3),
    CS_VR3F(3112062, SHAPE3_STAR, 1, 6, 2, 2, 4, 3),  CS_VR3F(3012062, SHAPE3_BOX, 1, 6, 2, 2, 4, 3),
    CS_VR3D(902023, SHAPE3_BOX, 2, 2, 3, 3),  CS_VR3D(904023, SHAPE3_BOX, 4, 2, 3, 2),
    CS_VR3D(902043, SHAPE3_BOX, 2, 4, 3, 2),  CS_VR3D(912023, SHAPE3_STAR, 2, 2, 3, 4),
    CS_VR3D(914023, SHAPE3_STAR, 4, 2, 3, 2), CS_VR3D(912043, SHAPE3_STAR, 2, 4, 3, 2),
};

const Variant *variants_3d_reg(int *count) {
  *count = (int)(sizeof(kVariantsReg3D) / sizeof(kVariantsReg3D[0]));
  return kVariantsReg3D;
}

}  // namespace cs
