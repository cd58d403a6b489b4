// k2d_reg.hip -- register-streaming FP64 2-D kernels for gfx950 (box and star, radius 1-3).
//
// Replaces kernel2d (reference src/2d/gpu.cu:30-70) and its time loop (src/2d/gpu.cu:316-329,
// 336-425) where vector FMAs beat the MFMA tables.  On gfx950 FP64 vector FMAs and FP64 MFMAs
// share one peak (128 flops per clock and CU); the dilated 8 x 8 MFMA tables execute about
// twice the useful multiply-adds of a 9-point box and more for the stars, while vector FMAs
// execute exactly the stencil's taps.
//
// A wave owns a strip of 64 PX columns (lane l: columns l PX .. l PX + PX - 1) and sweeps a
// chunk of rows downwards, one input row per iteration, advancing every row KF steps on
// the way.  Level s + 1 of the time loop keeps partial sums for the 2R + 1 rows that the
// incoming level-s row y contributes to: its dy = +R taps complete row y - R (which then
// becomes the incoming row of level s + 2), its dy = -R taps start row y + R, and the taps
// in between add to the rows in between.  Per level 2R partial rows persist between
// iterations; the incoming rows themselves are never stored.  The column neighbours of a
// lane's run come from the lanes on either side (v_mov_b32_dpp wave shifts, repeated for
// cells more than PX columns away; the lanes at the wave's ends receive zeros), so the
// valid columns shrink by R per level from each end of the strip: the central 64 PX - 2 R KF
// columns are the outputs and strips overlap by 2 R KF recomputed columns.  A chunk starts
// R KF rows above its first output row and ends R KF rows below its last.  Cells outside
// the interior keep their value at every level (the reference never writes the halo): when
// such a row or column arrives at a level its value is taken as the level's result and no
// later tap changes it.  No LDS, no barrier.
//
// Sum order per output: rows dy = -R .. R, columns dx = -R .. R within each (the
// reference's order), as a chain of FMAs.
#include "cs_internal.h"

namespace cs {

template <int R, int PX, int KF>
struct Reg2D {
  static constexpr int H = R * KF;            // recomputed columns per side; warm-up rows
  static constexpr int OX = 64 * PX - 2 * H;  // output columns per wave
  static constexpr int NS = 2 * R + 1;        // partial-sum rows per level
  static_assert(OX > 0, "strip too narrow for the fusion depth");
};

constexpr int modp(int v, int m) { return ((v % m) + m) % m; }

template <int R, int SH, int PX, int KF, int RING, bool EDGE>
struct Sweep2D {
  using C = Reg2D<R, PX, KF>;
  static constexpr int D = 2 * R + 1, NS = C::NS;
  static constexpr bool BOX = SH == SHAPE_BOX;
  static constexpr int PER = reg_period(R, RING);
  static constexpr int NL = (R + PX - 1) / PX;  // lanes a halo reaches across
  // The box radius-3 kernels hold more partial sums than the register file and spill; they
  // keep plain conditional stores, zero-selected checked loads and a prologue in free
  // order, which measured faster for them (box2d3r 8192^2, depth 2: 348 GStencil/s against
  // 102-182 with the buffer stores, clamped loads and ordered prologue of the others)
  static constexpr bool SPILLS = R > 1 && BOX;
  // The radius-3 star runs its first period ahead of the loop (see run); measured at 8192^2
  // 1018 GStencil/s against 982.  For radius 1 with 4 columns per lane only the plain path
  // does (box2d1r 10240^2: 2021-2038 against 1974-1977); on both paths the scheduling
  // barriers this needs cost registers (31 spilled, 1831), and 2 columns per lane gain
  // nothing.
  static constexpr bool PEEL = (R > 1 && !BOX) || (!EDGE && R == 1 && PX == 4);

  const double *__restrict__ in;
  double *__restrict__ out;
  double w[D * D];
  int64_t ny, nx, pitch, ey, oy, ox, cx, xs;
  int lane;
  int dcol[PX];  // checked path: column j + dcol[j] of the lane's run is column j clamped into
                 // the array (|dcol| <= 64 PX: a strip ends within 64 PX columns of the interior)
  bool colin[PX], colout[PX];
  double ring[RING][PX];
  double acc[KF][NS][PX];

  __device__ __forceinline__ void load_row(int64_t r, double (&dst)[PX]) {
    const int64_t ay = oy + r;
    if constexpr (!EDGE) {
      // every row and column the sweep touches is interior, hence inside the array
      const double *p = in + ay * pitch + ox + cx;
#pragma unroll
      for (int j = 0; j < PX; ++j) dst[j] = p[j];
    } else if constexpr (SPILLS) {
      const bool rowok = ay >= 0 && ay < ey;  // wave-uniform
      const double *p = in + (rowok ? ay : 0) * pitch + ox + cx;
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        const bool ok = rowok && dcol[j] == 0;
        const double v = *(ok ? p + j : in);
        dst[j] = ok ? v : 0.0;
      }
    } else {
      // row and columns clamped into the array: a cell outside it lies more than R cells
      // from the interior (the halo is at least R wide) and outside the interior every
      // cell is frozen, so what it holds reaches no stored output -- and no select waits
      // on the load
      const int64_t cy = ay < 0 ? 0 : ay >= ey ? ey - 1 : ay;  // wave-uniform
      const double *p = in + cy * pitch + ox + cx;
#pragma unroll
      for (int j = 0; j < PX; ++j) dst[j] = p[j + dcol[j]];
    }
  }

  // Output row yo of level KF from q, by buffer stores over the strip's row: a row outside
  // [y0, y1) gets an empty range and a lane's cells outside the strip's outputs an offset
  // past it, so every lane issues the same stores on every row (no branch in the loop).
  // The plain path stores a lane's cells in pairs (with H even both cells of a pair are
  // outputs or neither).
  __device__ __forceinline__ void store_row(int64_t yo, bool ok, const double (&q)[PX]) {
    if constexpr (SPILLS) {
      if (ok) {
        double *p = out + (oy + yo) * pitch + ox + cx;
#pragma unroll
        for (int j = 0; j < PX; ++j)
          if (colout[j]) p[j] = q[j];
      }
      return;
    }
    const __amdgpu_buffer_rsrc_t rs = buffer_over(out + (oy + yo) * pitch + ox + xs, ok ? 64 * PX * 8 : 0);
    const int base = lane * PX * 8;
    if constexpr (!EDGE && C::H % 2 == 0 && PX % 2 == 0) {
#pragma unroll
      for (int j = 0; j < PX; j += 2) buffer_store2(rs, colout[j] ? base + j * 8 : kBufferDrop, q[j], q[j + 1]);
    } else {
#pragma unroll
      for (int j = 0; j < PX; ++j) buffer_store1(rs, colout[j] ? base + j * 8 : kBufferDrop, q[j]);
    }
  }

  // qe[R + c] = cell c of the lane's run, c = -R .. PX + R - 1: the run itself and the R cells
  // on either side, which lie in the lanes up to NL away; lo / hi hold, after pass a, the
  // cells of lanes l - a / l + a that are still needed
  __device__ __forceinline__ void extend(const double (&q)[PX], double (&qe)[PX + 2 * R]) {
    double lo[PX], hi[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) {
      qe[R + j] = q[j];
      lo[j] = q[j];
      hi[j] = q[j];
    }
#pragma unroll
    for (int a = 1; a <= NL; ++a) {
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        if (j >= a * PX - R) {  // cell j - a PX in [-R, -1]
          lo[j] = lane_from_lower(lo[j]);
          qe[R + j - a * PX] = lo[j];
        }
        if (j < PX + R - a * PX) {  // cell j + a PX in [PX, PX + R - 1]
          hi[j] = lane_from_upper(hi[j]);
          qe[R + j + a * PX] = hi[j];
        }
      }
    }
  }

  // Iteration `it + K` (it a multiple of the period): input row r enters level 1, every
  // level passes its completed row on, level KF's completed row r - R KF is stored.  The
  // ring holds rows r .. r + RING - 2 (in flight); row r + RING - 1 is loaded into the
  // entry whose row r - 1 the previous iteration consumed.
  template <int K>
  __device__ __forceinline__ void phase(int64_t it, int64_t ys, int64_t y0, int64_t y1) {
    const int64_t r = ys + it + K;
    load_row(r + RING - 1, ring[(K + RING - 1) % RING]);
    double q[PX];
#pragma unroll
    for (int j = 0; j < PX; ++j) q[j] = ring[K % RING][j];
#pragma unroll
    for (int s = 0; s < KF; ++s) {
      // incoming: level s, row yq; level s + 1's partial sum for row y sits in slot
      // (y - ys) mod NS, and yq - ys = it + K - R s with `it` a multiple of NS
      const int64_t yq = r - R * s;
      double qe[PX + 2 * R];
      extend(q, qe);
#pragma unroll
      for (int dy = R; dy >= -R; --dy) {
        // contribution of the taps of row offset dy to row yq - dy (dy = R completes it,
        // dy = -R starts it)
        double(&T)[PX] = acc[s][modp(K - R * s - dy, NS)];
        const int64_t yt = yq - dy;
        const bool frozen = EDGE && (yt < 0 || yt >= ny);
#pragma unroll
        for (int j = 0; j < PX; ++j) {
          double t = dy == -R ? 0.0 : T[j];
          bool first = dy == -R;
#pragma unroll
          for (int dx = -R; dx <= R; ++dx) {
            if (!BOX && dy != 0 && dx != 0) continue;
            const double wv = w[(dy + R) * D + dx + R], v = qe[R + j + dx];
            t = first ? wv * v : __builtin_fma(wv, v, t);
            first = false;
          }
          if constexpr (EDGE) {
            // rows above yq arrived (and froze) earlier; row yq freezes now; rows below are
            // overwritten when they arrive
            if (dy > 0)
              t = frozen || !colin[j] ? T[j] : t;
            else if (dy == 0)
              t = frozen || !colin[j] ? q[j] : t;
          }
          T[j] = t;
        }
        // keep the partial sums computed here: left to itself the compiler sinks these FMAs
        // to their next use and keeps every level's incoming row alive instead
        if (dy < R) {
#pragma unroll
          for (int j = 0; j < PX; ++j) asm volatile("" : "+v"(T[j]));
        }
      }
      const double(&A)[PX] = acc[s][modp(K - R * s - R, NS)];
#pragma unroll
      for (int j = 0; j < PX; ++j) q[j] = A[j];
    }
    // q: level KF, row r - R KF
    const int64_t yo = r - C::H;
    store_row(yo, yo >= y0 && yo < y1, q);
    // each phase's memory operations stay in their phase: the scheduler would otherwise
    // sink the first period's loads below its stores (they are used only in the loop)
    if constexpr (PEEL) __builtin_amdgcn_sched_barrier(0);
  }

  template <int K>
  __device__ __forceinline__ void phases(int64_t it, int64_t ys, int64_t y0, int64_t y1) {
    if constexpr (K < PER) {
      phase<K>(it, ys, y0, y1);
      phases<K + 1>(it, ys, y0, y1);
    }
  }

  __device__ __forceinline__ void run(int64_t y0, int64_t y1) {
    const int64_t ys = y0 - C::H;
    const int64_t iters = (y1 - y0) + 2 * C::H;
#pragma unroll
    for (int s = 0; s < KF; ++s)
#pragma unroll
      for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int j = 0; j < PX; ++j) acc[s][k][j] = 0.0;
    // The partial-sum slots repeat every NS iterations, the ring every RING.  The loop head's
    // wait is one instruction for the first entry and every later one.  After the prologue
    // fewer operations follow each ring row than in the steady state (no stores), so that
    // wait is sized for the first entry and drains more of the ring than the steady state
    // needs.  With PEEL the first period runs before the loop (every sweep has at least one:
    // iters >= 1) and the loop is entered in the steady state.  Otherwise the prologue's rows
    // are at least issued in order (a scheduling barrier after each; issued youngest-first,
    // the head's wait drained the whole ring once per period), each followed by a row of
    // dropped stores, the loop's own pattern (the compiler merges those and moves them ahead
    // of the loads; the form is kept as measured: 1979-1985 GStencil/s for box2d1r).
#pragma unroll
    for (int k = 0; k < RING - 1; ++k) {
      load_row(ys + k, ring[k]);
      if constexpr (!SPILLS && !PEEL) {
        const double none[PX] = {};
        store_row(ys, false, none);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    int64_t it0 = 0;
    if constexpr (PEEL) {
      phases<0>(0, ys, y0, y1);
      it0 = PER;
    }
    for (int64_t it = it0; it < iters; it += PER) phases<0>(it, ys, y0, y1);
  }
};

template <int R, int SH, int PX, int KF, int RING, bool EDGE>
__device__ __forceinline__ void sweep(const double *__restrict__ in, double *__restrict__ out, const double *wts,
                                      const Geo &g, int64_t xs, int64_t y0, int64_t y1, int lane) {
  using C = Reg2D<R, PX, KF>;
  Sweep2D<R, SH, PX, KF, RING, EDGE> S;
  S.in = in;
  S.out = out;
#pragma unroll
  for (int i = 0; i < (2 * R + 1) * (2 * R + 1); ++i) S.w[i] = wts[i];
  S.ny = g.n[1];
  S.nx = g.n[2];
  S.pitch = g.ext[2];
  S.ey = g.ext[1];
  S.oy = g.org[1];
  S.ox = g.org[2];
  S.xs = xs;
  S.lane = lane;
  S.cx = xs + (int64_t)lane * PX;
#pragma unroll
  for (int j = 0; j < PX; ++j) {
    const int64_t c = S.cx + j, d = (int64_t)lane * PX + j;
    S.colin[j] = c >= 0 && c < S.nx;
    S.dcol[j] = (int)(S.ox + c < 0 ? -(S.ox + c) : S.ox + c >= S.pitch ? S.pitch - 1 - (S.ox + c) : 0);
    S.colout[j] = d >= C::H && d < C::H + C::OX && c < S.nx;  // c >= 0 wherever d >= H
  }
  S.run(y0, y1);
}

template <int R, int SH, int PX, int KF, int RING, int NW, int WPE>
__global__ __launch_bounds__(NW * 64, WPE) void k2d_reg(const double *__restrict__ in, double *__restrict__ out,
                                                        const double *__restrict__ wts, Geo g, RegSplit sp,
                                                        RegMulti mm) {
  using C = Reg2D<R, PX, KF>;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * NW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int64_t strip, y0, y1;
  if (mm.items) {
    // one item of a multi-grid launch (wave-uniform loads)
    if (wid >= mm.nitems) return;
    const RegItem it = mm.items[wid];
    const RegGrid &G = mm.grids[it.grid];
    g = G.geo;
    in = G.buf[G.cur ^ mm.flip];
    out = G.buf[G.cur ^ mm.flip ^ 1];
    wts = G.wts;
    strip = it.strip;
    y0 = it.y0;
    y1 = it.y1;
  } else if (!reg_chunk(sp, g.n[1], wid, &strip, &y0, &y1)) {
    return;
  }
  const int64_t xs = strip * C::OX - C::H;  // interior column of the strip's first cell
  // the plain path: every row the sweep freezes or loads (rows y0 - R KF .. y1 + R KF +
  // period + RING - 3) and every column of the strip interior
  constexpr int least = reg_least(R, KF, RING);
  if (xs >= 0 && xs + 64 * PX <= g.n[2] && y0 >= least && y1 + least <= g.n[1])
    sweep<R, SH, PX, KF, RING, false>(in, out, wts, g, xs, y0, y1, lane);
  else
    sweep<R, SH, PX, KF, RING, true>(in, out, wts, g, xs, y0, y1, lane);
}

// Resident waves of one variant on this device (workgroups per CU x waves x CUs), set
// by the variant's prepare() before its first launch.
template <int R, int SH, int PX, int KF, int RING, int NW, int WPE>
struct RegSlots {
  static int64_t value;
};
template <int R, int SH, int PX, int KF, int RING, int NW, int WPE>
int64_t RegSlots<R, SH, PX, KF, RING, NW, WPE>::value = 0;

template <int R, int SH, int PX, int KF, int RING, int NW, int WPE>
static hipError_t prepare_reg2d() {
  int nb = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, (const void *)k2d_reg<R, SH, PX, KF, RING, NW, WPE>, NW * 64, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  RegSlots<R, SH, PX, KF, RING, NW, WPE>::value = (int64_t)(nb > 0 ? nb : 1) * NW * (cus > 0 ? cus : 1);
  return hipSuccess;
}

template <int R, int SH, int PX, int KF, int RING, int NW, int WPE>
static hipError_t launch_reg2d(const double *in, double *out, const double *wts, const Geo &g, hipStream_t s) {
  using C = Reg2D<R, PX, KF>;
  using S = RegSlots<R, SH, PX, KF, RING, NW, WPE>;
  const int64_t slots = S::value > 0 ? S::value : 2048;
  const RegSplit sp = reg_split(g.n[1], g.n[2], C::OX, C::H, slots, reg_edge_cost(PX, R), reg_least(R, KF, RING));
  const int64_t waves = reg_waves(sp);
  hipLaunchKernelGGL((k2d_reg<R, SH, PX, KF, RING, NW, WPE>), dim3((unsigned)((waves + NW - 1) / NW)), dim3(NW * 64), 0,
                     s, in, out, wts, g, sp, RegMulti{});
  return hipGetLastError();
}

template <int R, int SH, int PX, int KF, int RING, int NW, int WPE>
static hipError_t launch_multi_reg2d(const RegMulti &mm, hipStream_t s) {
  if (mm.nitems <= 0) return hipSuccess;
  hipLaunchKernelGGL((k2d_reg<R, SH, PX, KF, RING, NW, WPE>), dim3((unsigned)((mm.nitems + NW - 1) / NW)),
                     dim3(NW * 64), 0, s, nullptr, nullptr, nullptr, Geo{}, RegSplit{}, mm);
  return hipGetLastError();
}

// Variant::mf 7 marks this family: by = 1 (rows are chunked per launch), bx = output
// columns per wave, px = columns per lane, st = rows in the load ring; WPE = the waves per
// SIMD the compiler is held to (registers); the weights are the dense (2R+1)^2 block.
#define CS_VR2D(ID, R, SH, PX, KF, RING, WPE)                                                                 \
  Variant {                                                                                                   \
    ID, 7, R, SH, KF, 1, Reg2D<R, PX, KF>::OX, 4, 0, PX, RING, &launch_reg2d<R, SH, PX, KF, RING, 4, WPE>,    \
        &prepare_reg2d<R, SH, PX, KF, RING, 4, WPE>,                                                         \
        "k2d_reg<r" #R "," #SH ",p" #PX ",kf" #KF ",ring" #RING ",wpe" #WPE ">", 0,                         \
        &launch_multi_reg2d<R, SH, PX, KF, RING, 4, WPE>                                                     \
  }

// ids 700000 + (ring of 2 (2R + 1) rows) * 50000 + (star) * 10000 + (R > 1) * R * 1000 + PX * 100
// + KF; 780408: the box radius-1 kernel with a ring of 4 rows, the default (launches of
// 0.400-0.414 ms against 0.417 for a ring of 3, box2d1r 2012-2074 GStencil/s; a ring of 2
// 1982).  Measured (GStencil/s over 64-step runs; boxes differ by up to 10%): box2d1r 10240^2
// p2 kf8 ring 6 (3 waves per SIMD) 1868-1940, p4 kf8 ring 3 (2 waves) 1950-2038, p2 kf12
// ring 6 (2 waves) 1849-1908, p2 kf10 ring 6 1840, p2 kf8 ring 3 1812-1830, p3 kf8 1715, p2
// kf6 ring 3 (4 waves) 1546; star2d3r 8192^2 p2 kf4 ring 7 1016-1018 (kf3 797, kf6 761);
// box2d3r p2 kf3 195, kf2 184 (these spill; the MFMA tables run it at 304-308).  Fewer
// columns per lane cost more recomputed columns and lane shifts per cell but leave
// registers for more waves and a deeper load ring, which the FP64 FMA issue needs: one wave
// issues an FP64 FMA about every 12 cycles, two waves about every 6.
static const Variant kVariantsReg2D[] = {
    CS_VR2D(780408, 1, SHAPE_BOX, 4, 8, 4, 2),   CS_VR2D(700408, 1, SHAPE_BOX, 4, 8, 3, 2),
    CS_VR2D(750208, 1, SHAPE_BOX, 2, 8, 6, 3),
    CS_VR2D(750212, 1, SHAPE_BOX, 2, 12, 6, 2),  CS_VR2D(700212, 1, SHAPE_BOX, 2, 12, 3, 3),
    CS_VR2D(750216, 1, SHAPE_BOX, 2, 16, 6, 2),  CS_VR2D(700404, 1, SHAPE_BOX, 4, 4, 3, 2),
    CS_VR2D(700402, 1, SHAPE_BOX, 4, 2, 3, 2),   CS_VR2D(700401, 1, SHAPE_BOX, 4, 1, 3, 2),
    CS_VR2D(710212, 1, SHAPE_STAR, 2, 12, 3, 3), CS_VR2D(760212, 1, SHAPE_STAR, 2, 12, 6, 2),
    CS_VR2D(760216, 1, SHAPE_STAR, 2, 16, 6, 2), CS_VR2D(760208, 1, SHAPE_STAR, 2, 8, 6, 3),
    CS_VR2D(710404, 1, SHAPE_STAR, 4, 4, 3, 2),  CS_VR2D(710402, 1, SHAPE_STAR, 4, 2, 3, 2),
    CS_VR2D(710401, 1, SHAPE_STAR, 4, 1, 3, 2),
    // radius 3 (star2d3r, box2d3r): rings of 7 rows
    CS_VR2D(713204, 3, SHAPE_STAR, 2, 4, 7, 2),  CS_VR2D(713203, 3, SHAPE_STAR, 2, 3, 7, 3),
    CS_VR2D(713206, 3, SHAPE_STAR, 2, 6, 7, 2),  CS_VR2D(713202, 3, SHAPE_STAR, 2, 2, 7, 4),
    CS_VR2D(713201, 3, SHAPE_STAR, 2, 1, 7, 4),  CS_VR2D(703203, 3, SHAPE_BOX, 2, 3, 7, 3),
    CS_VR2D(703202, 3, SHAPE_BOX, 2, 2, 7, 4),   CS_VR2D(703201, 3, SHAPE_BOX, 2, 1, 7, 4),
};

const Variant *variants_2d_reg(int *count) {
  *count = (int)(sizeof(kVariantsReg2D) / sizeof(kVariantsReg2D[0]));
  return kVariantsReg2D;
}

}  // namespace cs
