// k2d_stream.hip -- pipelined persistent FP64 2-D kernel: tile loads overlap the MFMA steps.
//
// Same tile computation as k2d_f64 (reference kernel2d, src/2d/gpu.cu:30-70, with the
// time loop of gpu_box_2d1r fused inside, src/2d/gpu.cu:316-329): dilated
// v_mfma_f64_4x4x4 blocks, KF fused steps per tile in LDS (k2d_f64_body.h).  What
// changes is how HBM and the matrix cores overlap.  In the one-tile-per-workgroup kernel
// every workgroup loads its tile, then computes, then stores; the workgroups of a launch
// start together, so the whole chip alternates between an HBM phase and an MFMA phase,
// and a launch takes about the sum of the two.  Here:
//   * one persistent workgroup per CU walks its tiles (XCD-aware order) with two LDS tile
//     buffers: while the consumer waves advance tile k in one buffer, the producer waves
//     have tile k + 1 in flight into the other;
//   * interior tiles are copied global -> LDS directly (global_load_lds_dword): the
//     loads hold no registers and the consumer waves never wait for them; only the
//     producer waves' own vmcnt tracks them, and they publish the tile at the
//     workgroup barrier that ends tile k.  Tiles at the domain edge are zero-filled
//     first and then copied the same way, with the lanes outside the array masked off;
//   * the producers take part in every barrier of the consumers' fused steps without
//     doing anything else there, so the consumers' code is that of k2d_f64 unchanged;
//   * the output block is staged in LDS and written row by row by the consumers (ST 2),
//     or written from the blocks' registers (ST 1).
#include "k2d_f64_body.h"

namespace cs {

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// Rows [r0, r1) of a tile (EY x EX cells from (ay0, ax0)) into LDS rows of pitch PX,
// columns from CS on: each row is ceil(2 EX / 64) 256-byte chunks, one
// global_load_lds_dword per chunk (lane i copies dword i of the chunk), rows dealt
// round-robin to the NP producer waves.  Nothing here waits for the copies; the caller
// waits for vmcnt(0).  In a tile at the domain edge only the cells inside the array are
// copied (rows and lanes outside are masked off); the rest must have been zeroed first.
template <class C, int NP>
__device__ __forceinline__ void dma_rows(const double *__restrict__ in, const Geo &g, const TileCtx &t, double *buf,
                                         int pw, int lane, int r0, int r1) {
  constexpr int DW = 2 * C::EX;            // dwords per row
  constexpr int CH = (DW + 63) / 64;       // chunks per row
  constexpr int TAIL = DW - 64 * (CH - 1);  // active lanes of the last chunk
  const char *g0 = reinterpret_cast<const char *>(in + t.ay0 * t.pitch + t.ax0) + lane * 4;
  char *l0 = reinterpret_cast<char *>(buf + C::CS);
  const int64_t rstep = t.pitch * 8;
  for (int r = r0 + pw; r < r1; r += NP) {
    const int64_t ay = t.ay0 + r;
    if (!t.full && (ay < 0 || ay >= g.ext[1])) continue;  // uniform over the wave
    const char *gr = g0 + r * rstep;
    char *l = l0 + r * C::PX * 8;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      bool on = c < CH - 1 || lane < TAIL;
      if (!t.full) {
        const int64_t ax = t.ax0 + ((64 * c + lane) >> 1);
        on = on && ax >= 0 && ax < t.pitch;
      }
      if (on) __builtin_amdgcn_global_load_lds((glb_void_t *)(gr + c * 256), (lds_void_t *)(l + c * 256), 4, 0, 0);
    }
  }
}

// A tile at the domain edge starts from zeros (ds_write, completed before any copy).
template <class C, int NP>
__device__ __forceinline__ void zero_tile(const TileCtx &t, double *buf, int ptid) {
  if (t.full) return;
  for (int i = ptid; i < C::BUF; i += NP * 64) buf[i] = 0.0;
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// Position k (step G) of the XCD-ordered tile walk, advanced to the next existing tile.
__device__ __forceinline__ int64_t stream_next(int64_t k, int64_t G, int64_t total, int64_t ntiles) {
  while (k < total && xcd_tile(k, ntiles) >= ntiles) k += G;
  return k;
}

template <int R, int SH, int KF, int BY, int BX, int NW, int NP, int ST>
__global__ __launch_bounds__((NW + NP) * 64) void k2d_stream(const double *__restrict__ in, double *__restrict__ out,
                                                            const double *__restrict__ bfrag, Geo g) {
  using C = Cfg2D<4, R, SH, KF, BY, BX, NW, ST>;
  static_assert(C::NBUF == 1, "one LDS buffer per tile");
  constexpr int NBAR = ST == 2 ? 2 * KF : 2 * (KF - 1);  // barriers inside one tile's steps
  constexpr int SLOT = C::BUF + 8;                        // doubles per tile buffer (+ spare cells)
  // the next tile is issued in NPH slices of RPH rows, one before each of the first NPH
  // step barriers (the last two steps' barriers get none, so the copies land early)
  constexpr int NPH = NBAR > 2 ? NBAR - 2 : 1;
  constexpr int RPH = (C::EY + NPH - 1) / NPH;
  extern __shared__ __align__(16) double lds[];
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t total = ((ntiles + 7) / 8) * 8, G = gridDim.x;
  int64_t k = stream_next(blockIdx.x, G, total, ntiles);
  if (k >= total) return;  // uniform over the workgroup

  // pad columns and spare cells are read by blocks whose results are never used; they must
  // hold finite values (0 * NaN would be NaN), so both buffers start at zero
  for (int i = threadIdx.x; i < 2 * SLOT; i += (NW + NP) * 64) lds[i] = 0.0;
  __syncthreads();

  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  int cur = 0;
  if (wave >= NW) {
    // ---- producers ----
    const int pw = wave - NW, lane = threadIdx.x & 63, ptid = threadIdx.x - NW * 64;
    TileCtx t;
    t.orow = t.ocol = 0;
    {
      tile_setup<C>(g, xcd_tile(k, ntiles), t, out);
      zero_tile<C, NP>(t, lds, ptid);
      dma_rows<C, NP>(in, g, t, lds, pw, lane, 0, C::EY);
      asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
      __builtin_amdgcn_s_barrier();  // tile k is in buffer 0
    }
    for (;;) {
      const int64_t kn = stream_next(k + G, G, total, ntiles);
      double *nb = lds + (cur ^ 1) * SLOT;
      const bool more = kn < total;
      if (more) {
        tile_setup<C>(g, xcd_tile(kn, ntiles), t, out);
        zero_tile<C, NP>(t, nb, ptid);
      }
      // the next tile's rows, a slice before each of the consumers' step barriers of tile k
      // (a wave may have at most 63 memory instructions in flight: issuing a whole tile at
      // once would stall the producers, and with them the consumers at their first barrier)
#pragma unroll
      for (int ph = 0; ph < NBAR; ++ph) {
        if (more && ph < NPH) dma_rows<C, NP>(in, g, t, nb, pw, lane, ph * RPH, min(C::EY, (ph + 1) * RPH));
        __builtin_amdgcn_s_barrier();
      }
      if (more && NBAR < NPH) dma_rows<C, NP>(in, g, t, nb, pw, lane, NBAR * RPH, C::EY);
      asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory");
      __builtin_amdgcn_s_barrier();  // end of tile k: tile kn is in the other buffer
      if (kn >= total) break;
      k = kn;
      cur ^= 1;
    }
  } else {
    // ---- consumers: the fused steps of k2d_f64 on the tile in buffer `cur` ----
    TileCtx t;
    double b[C::NKS];
    int off[C::NOFF];
    lane_setup<C>(bfrag, threadIdx.x, t, b, off);
    __builtin_amdgcn_s_barrier();  // the first tile is in buffer 0
    for (;;) {
      // opaque copies of the thread and wave indices: keep the compiler from hoisting the
      // per-tile addresses of every block out of the tile loop
      int tid = t.tid, w = t.wave;
      asm volatile("" : "+v"(tid), "+s"(w));
      t.tid = tid;
      t.wave = w;
      tile_setup<C>(g, xcd_tile(k, ntiles), t, out);
      char *buf = reinterpret_cast<char *>(lds + cur * SLOT);
      if (t.full)
        steps4<C, 0, false>(t, b, off, buf);
      else
        steps4<C, 0, true>(t, b, off, buf);
      lds_barrier();  // end of tile k: every read of this buffer is done
      const int64_t kn = stream_next(k + G, G, total, ntiles);
      if (kn >= total) break;
      k = kn;
      cur ^= 1;
    }
  }
}

template <int R, int SH, int KF, int BY, int BX, int NW, int NP, int ST>
struct StreamCfg {
  using C = Cfg2D<4, R, SH, KF, BY, BX, NW, ST>;
  static constexpr size_t LDS = (size_t)2 * (C::BUF + 8) * sizeof(double);
  static_assert(LDS <= 160 * 1024, "two tile buffers do not fit gfx950 LDS");
  static_assert((NW + NP) * 64 <= 1024, "workgroup too large");
};

template <int R, int SH, int KF, int BY, int BX, int NW, int NP, int ST>
static int &stream_grid() {
  static int grid = 0;
  return grid;
}

template <int R, int SH, int KF, int BY, int BX, int NW, int NP, int ST>
static hipError_t prepare_stream() {
  using S = StreamCfg<R, SH, KF, BY, BX, NW, NP, ST>;
  const void *fn = (const void *)k2d_stream<R, SH, KF, BY, BX, NW, NP, ST>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::LDS);
  if (e != hipSuccess) return e;
  int per_cu = 0, dev = 0, cus = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (NW + NP) * 64, S::LDS);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  stream_grid<R, SH, KF, BY, BX, NW, NP, ST>() = ((per_cu > 0 ? per_cu : 1) * cus) / 8 * 8;
  return hipSuccess;
}

template <int R, int SH, int KF, int BY, int BX, int NW, int NP, int ST>
static hipError_t launch_stream(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using S = StreamCfg<R, SH, KF, BY, BX, NW, NP, ST>;
  const int64_t total = ((g.tiles[1] * g.tiles[2] + 7) / 8) * 8;
  int64_t grid = stream_grid<R, SH, KF, BY, BX, NW, NP, ST>();
  if (grid <= 0) return hipErrorNotInitialized;  // prepare() sizes the grid
  if (grid > total) grid = total;
  hipLaunchKernelGGL((k2d_stream<R, SH, KF, BY, BX, NW, NP, ST>), dim3((unsigned)grid), dim3((NW + NP) * 64), S::LDS,
                     s, in, out, bfrag, g);
  return hipGetLastError();
}

#define CS_V2DS(ID, R, SH, KF, BY, BX, NW, NP, ST)                                                         \
  Variant {                                                                                                \
    ID, 4, R, SH, KF, BY, BX, NW, StreamCfg<R, SH, KF, BY, BX, NW, NP, ST>::LDS,                            \
        Cfg2D<4, R, SH, KF, BY, BX, NW, ST>::PX, ST, &launch_stream<R, SH, KF, BY, BX, NW, NP, ST>,        \
        &prepare_stream<R, SH, KF, BY, BX, NW, NP, ST>,                                                    \
        "k2d_stream<r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ",p" #NP ",st" #ST ">",              \
        Cfg2D<4, R, SH, KF, BY, BX, NW, ST>::CS                                                             \
  }

// Variant ids 800000 + R * 1000 + (SH == star) * 10 + KF % 10 + 100 * (tile code): tile code 1
// = 64 x 64 with 8 consumer waves, 2 = 96 x 64 / 8, 3 = 64 x 64 / 12 (output from registers),
// 4 = 80 x 80 / 10; two producer waves, or four for codes 5 (64 x 64 / 8) and 6 (96 x 64 / 8).
static const Variant kStreamVariants[] = {
    CS_V2DS(811106, 1, SHAPE_BOX, 6, 64, 64, 8, 2, 2),   CS_V2DS(811104, 1, SHAPE_BOX, 4, 64, 64, 8, 2, 2),
    CS_V2DS(811108, 1, SHAPE_BOX, 8, 64, 64, 8, 2, 2),   CS_V2DS(811206, 1, SHAPE_BOX, 6, 96, 64, 8, 2, 2),
    CS_V2DS(811306, 1, SHAPE_BOX, 6, 64, 64, 12, 2, 1),  CS_V2DS(811406, 1, SHAPE_BOX, 6, 80, 80, 10, 2, 2),
    CS_V2DS(811116, 1, SHAPE_STAR, 6, 64, 64, 8, 2, 2),  CS_V2DS(813103, 3, SHAPE_BOX, 3, 64, 64, 8, 2, 2),
    CS_V2DS(811506, 1, SHAPE_BOX, 6, 64, 64, 8, 4, 2),   CS_V2DS(811606, 1, SHAPE_BOX, 6, 96, 64, 8, 4, 2),
    CS_V2DS(813113, 3, SHAPE_STAR, 3, 64, 64, 8, 2, 2),
};

const Variant *variants_2d_stream(int *count) {
  *count = (int)(sizeof(kStreamVariants) / sizeof(kStreamVariants[0]));
  return kStreamVariants;
}

}  // namespace cs
