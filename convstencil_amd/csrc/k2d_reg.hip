// k2d_reg.hip -- register-streaming FP64 2-D kernels for gfx950 (radius 1, box and star).
//
// Replaces kernel2d (reference src/2d/gpu.cu:30-70) for the radius-1 time loop
// (src/2d/gpu.cu:316-329).  On gfx950 FP64 vector FMAs and FP64 MFMAs share one peak
// (128 flops per clock and CU); the dilated 8 x 8 MFMA tables execute about twice the
// useful multiply-adds of a 9-point stencil, while vector FMAs execute exactly those.
//
// A wave owns a strip of 64 PX columns (lane l: columns l PX .. l PX + PX - 1) and sweeps a
// chunk of rows downwards, one input row per iteration, advancing every row KF steps on
// the way.  Level s + 1 of the time loop keeps partial sums for the rows that the
// incoming level-s row contributes to: an incoming row y adds its dy = +1 taps to row
// y - 1 (which is then complete and becomes the incoming row of level s + 2), its dy = 0
// taps to row y and starts row y + 1 with its dy = -1 taps.  Per level two partial rows
// persist between iterations; the incoming rows themselves are never stored.  The column
// neighbours of a lane's run come from lanes l -+ 1 (v_mov_b32_dpp wave shifts; lanes 0
// and 63 receive zeros), so the valid columns shrink by one per level from each end of
// the strip: the central 64 PX - 2 KF columns are the outputs and strips overlap by 2 KF
// recomputed columns.  A chunk starts KF rows above its first output row and ends KF
// rows below its last.  Cells outside the interior keep their value at every level (the
// reference never writes the halo): when such a row or column arrives at a level its
// value is taken as the level's result and no later tap changes it.  No LDS, no barrier.
//
// Sum order per output: rows dy = -1, 0, +1, columns dx = -1, 0, +1 within each (the
// reference's order), as a chain of FMAs.
#include "cs_internal.h"

namespace cs {

template <int PX, int KF>
struct Reg2D {
  static constexpr int H = KF;                // recomputed columns per side; warm-up rows
  static constexpr int OX = 64 * PX - 2 * H;  // output columns per wave
  static_assert(OX > 0, "strip too narrow for the fusion depth");
};


constexpr int mod3(int v) { return ((v % 3) + 3) % 3; }


template <int SH, int PX, int KF, int RING, bool EDGE>
struct Sweep2D {
  using C = Reg2D<PX, KF>;
  static constexpr bool BOX = SH == SHAPE_BOX;

  const double *__restrict__ in;
  double *__restrict__ out;
  double w[9];
  int64_t ny, nx, pitch, ey, oy, ox, cx;
  bool colin[PX], colarr[PX], colout[PX];
  double ring[RING][PX];
  double acc[KF][3][PX];

  __device__ __forceinline__ void load_row(int64_t r, double (&dst)[PX]) {
    const int64_t ay = oy + r;
    if constexpr (!EDGE) {
      // every row and column the sweep touches is interior, hence inside the array
      const double *p = in + ay * pitch + ox + cx;
#pragma unroll
      for (int j = 0; j < PX; ++j) dst[j] = p[j];
    } else {
      const bool rowok = ay >= 0 && ay < ey;  // wave-uniform
      const double *p = in + (rowok ? ay : 0) * pitch + ox + cx;
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        const bool ok = rowok && colarr[j];
        const double v = *(ok ? p + j : in);
        dst[j] = ok ? v : 0.0;
      }
    }
  }

  // Iteration `it + K` (it a multiple of the period): input row r enters level 1, every
  // level passes its completed row on, level KF's completed row r - KF is stored.  The
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
      // incoming: level s, row yq; accumulators of level s + 1 for rows yq - 1, yq, yq + 1
      const int64_t yq = r - s;
      const int sA = mod3(K % 3 - s - 1), sB = mod3(K % 3 - s), sC = mod3(K % 3 - s + 1);
      double(&A)[PX] = acc[s][sA];
      double(&B)[PX] = acc[s][sB];
      double(&Cn)[PX] = acc[s][sC];
      const double ql = lane_from_lower(q[PX - 1]), qr = lane_from_upper(q[0]);
      double qe[PX + 2];
      qe[0] = ql;
      qe[PX + 1] = qr;
#pragma unroll
      for (int j = 0; j < PX; ++j) qe[j + 1] = q[j];
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        double a = A[j], b = B[j];
        if constexpr (BOX) {
          a = __builtin_fma(w[6], qe[j], a);
          a = __builtin_fma(w[7], qe[j + 1], a);
          a = __builtin_fma(w[8], qe[j + 2], a);
        } else {
          a = __builtin_fma(w[7], qe[j + 1], a);
        }
        b = __builtin_fma(w[3], qe[j], b);
        b = __builtin_fma(w[4], qe[j + 1], b);
        b = __builtin_fma(w[5], qe[j + 2], b);
        if constexpr (EDGE) {
          // a cell outside the interior keeps the value it arrived with: row yq - 1 (or this
          // column) arrived frozen one iteration ago, row yq arrives now
          const bool fa = yq - 1 < 0 || yq - 1 >= ny || !colin[j];
          const bool fb = yq < 0 || yq >= ny || !colin[j];
          A[j] = fa ? A[j] : a;
          B[j] = fb ? q[j] : b;
        } else {
          A[j] = a;
          B[j] = b;
        }
      }
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        if constexpr (BOX) {
          double c = w[0] * qe[j];
          c = __builtin_fma(w[1], qe[j + 1], c);
          Cn[j] = __builtin_fma(w[2], qe[j + 2], c);
        } else {
          Cn[j] = w[1] * qe[j + 1];
        }
      }
      // keep the partial sums computed here: left to itself the compiler sinks these FMAs
      // to their next use and keeps every level's incoming row alive instead
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        asm volatile("" : "+v"(B[j]), "+v"(Cn[j]));
        q[j] = A[j];
      }
    }
    // q: level KF, row r - KF
    const int64_t yo = r - KF;
    if (yo >= y0 && yo < y1) {
      double *p = out + (oy + yo) * pitch + ox + cx;
#pragma unroll
      for (int j = 0; j < PX; ++j)
        if (colout[j]) p[j] = q[j];
    }
  }

  __device__ __forceinline__ void run(int64_t y0, int64_t y1) {
    const int64_t ys = y0 - C::H;
    const int64_t iters = (y1 - y0) + 2 * C::H;
#pragma unroll
    for (int s = 0; s < KF; ++s)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < PX; ++j) acc[s][k][j] = 0.0;
#pragma unroll
    for (int k = 0; k < RING - 1; ++k) load_row(ys + k, ring[k]);
    // the accumulator slots repeat every 3 iterations, the ring every RING
    for (int64_t it = 0; it < iters; it += reg_period(RING)) phases<0>(it, ys, y0, y1);
  }

  template <int K>
  __device__ __forceinline__ void phases(int64_t it, int64_t ys, int64_t y0, int64_t y1) {
    if constexpr (K < reg_period(RING)) {
      phase<K>(it, ys, y0, y1);
      phases<K + 1>(it, ys, y0, y1);
    }
  }
};

template <int SH, int PX, int KF, int RING, bool EDGE>
__device__ __forceinline__ void sweep(const double *__restrict__ in, double *__restrict__ out, const double *wts,
                                      const Geo &g, int64_t xs, int64_t y0, int64_t y1, int lane) {
  using C = Reg2D<PX, KF>;
  Sweep2D<SH, PX, KF, RING, EDGE> S;
  S.in = in;
  S.out = out;
#pragma unroll
  for (int i = 0; i < 9; ++i) S.w[i] = wts[i];
  S.ny = g.n[1];
  S.nx = g.n[2];
  S.pitch = g.ext[2];
  S.ey = g.ext[1];
  S.oy = g.org[1];
  S.ox = g.org[2];
  S.cx = xs + (int64_t)lane * PX;
#pragma unroll
  for (int j = 0; j < PX; ++j) {
    const int64_t c = S.cx + j, d = (int64_t)lane * PX + j;
    S.colin[j] = c >= 0 && c < S.nx;
    S.colarr[j] = S.ox + c >= 0 && S.ox + c < S.pitch;
    S.colout[j] = d >= C::H && d < C::H + C::OX && c < S.nx;  // c >= 0 wherever d >= H
  }
  S.run(y0, y1);
}

template <int SH, int PX, int KF, int RING, int NW, int WPE>
__global__ __launch_bounds__(NW * 64, WPE) void k2d_reg(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ wts, Geo g, RegSplit sp) {
  using C = Reg2D<PX, KF>;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * NW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int64_t strip, y0, y1;
  if (!reg_chunk(sp, g.n[1], wid, &strip, &y0, &y1)) return;
  const int64_t xs = strip * C::OX - C::H;  // interior column of the strip's first cell
  // the plain path: every row the sweep freezes or loads (rows y0 - KF .. y1 + KF + period +
  // RING - 3) and every column of the strip interior
  if (xs >= 0 && xs + 64 * PX <= g.n[2] && y0 >= reg_least(KF, RING) && y1 + reg_least(KF, RING) <= g.n[1])
    sweep<SH, PX, KF, RING, false>(in, out, wts, g, xs, y0, y1, lane);
  else
    sweep<SH, PX, KF, RING, true>(in, out, wts, g, xs, y0, y1, lane);
}

// Resident waves of one variant on this device (workgroups per CU x waves x CUs), set
// by the variant's prepare() before its first launch.
template <int SH, int PX, int KF, int RING, int NW, int WPE>
struct RegSlots {
  static int64_t value;
};
template <int SH, int PX, int KF, int RING, int NW, int WPE>
int64_t RegSlots<SH, PX, KF, RING, NW, WPE>::value = 0;

template <int SH, int PX, int KF, int RING, int NW, int WPE>
static hipError_t prepare_reg2d() {
  int nb = 0, dev = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, (const void *)k2d_reg<SH, PX, KF, RING, NW, WPE>, NW * 64, 0);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  RegSlots<SH, PX, KF, RING, NW, WPE>::value = (int64_t)(nb > 0 ? nb : 1) * NW * (cus > 0 ? cus : 1);
  return hipSuccess;
}

template <int SH, int PX, int KF, int RING, int NW, int WPE>
static hipError_t launch_reg2d(const double *in, double *out, const double *wts, const Geo &g, hipStream_t s) {
  using C = Reg2D<PX, KF>;
  using S = RegSlots<SH, PX, KF, RING, NW, WPE>;
  const int64_t slots = S::value > 0 ? S::value : 2048;
  const RegSplit sp = reg_split(g.n[1], g.n[2], C::OX, KF, slots, kRegEdgeCost, reg_least(KF, RING));
  const int64_t waves = reg_waves(sp);
  hipLaunchKernelGGL((k2d_reg<SH, PX, KF, RING, NW, WPE>), dim3((unsigned)((waves + NW - 1) / NW)), dim3(NW * 64), 0,
                     s, in, out, wts, g, sp);
  return hipGetLastError();
}

// Variant::mf 7 marks this family: by = 1 (rows are chunked per launch), bx = output
// columns per wave, px = columns per lane, st = rows in the load ring; WPE = the waves per
// SIMD the compiler is held to (registers); the weights are the dense 3 x 3 block.
#define CS_VR2D(ID, SH, PX, KF, RING, WPE)                                                                     \
  Variant {                                                                                                    \
    ID, 7, 1, SH, KF, 1, Reg2D<PX, KF>::OX, 4, 0, PX, RING, &launch_reg2d<SH, PX, KF, RING, 4, WPE>,          \
        &prepare_reg2d<SH, PX, KF, RING, 4, WPE>,                                                             \
        "k2d_reg<r1," #SH ",p" #PX ",kf" #KF ",ring" #RING ",wpe" #WPE ">"                                   \
  }

// ids 700000 + (ring 6) * 50000 + (star) * 10000 + PX * 100 + KF.  Measured at 10240^2
// (GStencil/s over 64-step runs, one box): box p2 kf12 ring 6 (2 waves per SIMD) 1902,
// p2 kf12 ring 3 (3 waves) 1827, p2 kf8 ring 6 (3 waves) 1722, p2 kf16 ring 6 1659, p4 kf8
// ring 3 (2 waves) 1428; star p2 kf12 ring 3 2428.  Fewer columns per lane cost more
// recomputed columns and lane shifts per cell but leave registers for more waves and a
// deeper load ring, which the FP64 FMA issue needs: one wave issues an FP64 FMA about every
// 12 cycles, two waves about every 6.
static const Variant kVariantsReg2D[] = {
    CS_VR2D(750212, SHAPE_BOX, 2, 12, 6, 2),  CS_VR2D(700212, SHAPE_BOX, 2, 12, 3, 3),
    CS_VR2D(750208, SHAPE_BOX, 2, 8, 6, 3),   CS_VR2D(750216, SHAPE_BOX, 2, 16, 6, 2),
    CS_VR2D(700408, SHAPE_BOX, 4, 8, 3, 2),   CS_VR2D(700404, SHAPE_BOX, 4, 4, 3, 2),
    CS_VR2D(700402, SHAPE_BOX, 4, 2, 3, 2),   CS_VR2D(700401, SHAPE_BOX, 4, 1, 3, 2),
    CS_VR2D(710212, SHAPE_STAR, 2, 12, 3, 3), CS_VR2D(760212, SHAPE_STAR, 2, 12, 6, 2),
    CS_VR2D(760216, SHAPE_STAR, 2, 16, 6, 2), CS_VR2D(760208, SHAPE_STAR, 2, 8, 6, 3),
    CS_VR2D(710404, SHAPE_STAR, 4, 4, 3, 2),  CS_VR2D(710402, SHAPE_STAR, 4, 2, 3, 2),
    CS_VR2D(710401, SHAPE_STAR, 4, 1, 3, 2),
};

const Variant *variants_2d_reg(int *count) {
  *count = (int)(sizeof(kVariantsReg2D) / sizeof(kVariantsReg2D[0]));
  return kVariantsReg2D;
}

}  // namespace cs
