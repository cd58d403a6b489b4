// k2d_f64.hip -- plain FP64 2-D ConvStencil kernels (see k2d_f64_body.h for the tile body).
#include "k2d_f64_body.h"

namespace cs {

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
__global__ __launch_bounds__(NW * 64) void k2d_f64(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  extern __shared__ __align__(16) double lds[];
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;  // uniform over the workgroup
  tile_f64<MF, R, SH, KF, BY, BX, NW, ST>(in, out, bfrag, g, tile, lds, threadIdx.x);
}

// The same kernel with 128-byte aligned LDS rows (the --breakdown ladder's rung 3).
template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
__global__ __launch_bounds__(NW * 64) void k2d_f64u(const double *__restrict__ in, double *__restrict__ out,
                                                    const double *__restrict__ bfrag, Geo g) {
  extern __shared__ __align__(16) double lds[];
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;  // uniform over the workgroup
  tile_f64<MF, R, SH, KF, BY, BX, NW, ST, 64, kAligned>(in, out, bfrag, g, tile, lds, threadIdx.x);
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static hipError_t prepare_k2du() {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST, kAligned>;
  return hipFuncSetAttribute((const void *)k2d_f64u<MF, R, SH, KF, BY, BX, NW, ST>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static hipError_t launch_k2du(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST, kAligned>;
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k2d_f64u<MF, R, SH, KF, BY, BX, NW, ST>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in,
                     out, bfrag, g);
  return hipGetLastError();
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static hipError_t prepare_k2d() {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>;
  return hipFuncSetAttribute((const void *)k2d_f64<MF, R, SH, KF, BY, BX, NW, ST>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static hipError_t launch_k2d(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>;
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k2d_f64<MF, R, SH, KF, BY, BX, NW, ST>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in,
                     out, bfrag, g);
  return hipGetLastError();
}

// ---- persistent form: a fixed grid walks the tiles, prefetching the next one ----
//
// As many workgroups as fit on the device at once (a multiple of 8, so that each stays
// on one XCD's band of tiles) take tiles w, w + G, w + 2G, ... in the XCD-aware order.
// While a workgroup computes one tile, one load per thread touches every cache line of
// its next tile, so that tile's own load (at the start of its turn) comes from L2 rather
// than HBM; the fused steps synchronise through LDS only (lds_barrier), so nothing waits
// for those loads until the tile is done.

// Next position >= k (step G) whose XCD-ordered tile exists; returns total if none.
__device__ __forceinline__ int64_t next_tile_pos(int64_t k, int64_t G, int64_t total, int64_t ntiles) {
  while (k < total && xcd_tile(k, ntiles) >= ntiles) k += G;
  return k;
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
__global__ __launch_bounds__(NW * 64) void k2d_f64p(const double *__restrict__ in, double *__restrict__ out,
                                                    const double *__restrict__ bfrag, Geo g) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>;
  extern __shared__ __align__(16) double lds[];
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t total = ((ntiles + 7) / 8) * 8, G = gridDim.x;
  int64_t k = next_tile_pos(blockIdx.x, G, total, ntiles);
  if (k >= total) return;
  TileCtx t;
  double b[C::NKS];
  int off[C::NOFF];
  lane_setup<C>(bfrag, threadIdx.x, t, b, off);
  tile_setup<C>(g, xcd_tile(k, ntiles), t, out);
  for (;;) {
    // opaque copies of the thread and wave indices: keep the compiler from hoisting the
    // per-element load addresses and every block's addresses out of the tile loop
    // (they would all stay live in registers)
    int tid = t.tid, w = t.wave;
    asm volatile("" : "+v"(tid), "+s"(w));
    t.tid = tid;
    t.wave = w;
    if constexpr (C::MF == 4) {
      if (t.full)
        load_tile_full<C>(in, t, lds);
      else
        load_tile<C, 64>(in, g, t, lds);
    } else {
      load_tile<C, 64>(in, g, t, lds);
    }
    lds_barrier();
    const int64_t kn = next_tile_pos(k + G, G, total, ntiles);
    const bool more = kn < total;
    TileCtx tn = t;
    double pf[Prefetch<C>::N];
    if (more) {
      tile_setup<C>(g, xcd_tile(kn, ntiles), tn, out);
      prefetch_tile<C>(in, g, tn, pf);  // in flight while this tile is computed
    }
    run_steps<C>(t, b, off, lds, out);
    if (!more) break;
    prefetch_consume<C>(pf);
    lds_barrier();  // every wave is done with this tile's LDS before the next tile is stored
    k = kn;
    t = tn;
  }
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static int &k2dp_grid() {
  static int grid = 0;
  return grid;
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static hipError_t prepare_k2dp() {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>;
  const void *fn = (const void *)k2d_f64p<MF, R, SH, KF, BY, BX, NW, ST>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
  if (e != hipSuccess) return e;
  int per_cu = 0, dev = 0, cus = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NW * 64, C::LDS);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  k2dp_grid<MF, R, SH, KF, BY, BX, NW, ST>() = ((per_cu > 0 ? per_cu : 1) * cus) / 8 * 8;
  return hipSuccess;
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
static hipError_t launch_k2dp(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>;
  const int64_t total = ((g.tiles[1] * g.tiles[2] + 7) / 8) * 8;
  int64_t grid = k2dp_grid<MF, R, SH, KF, BY, BX, NW, ST>();
  if (grid <= 0) return hipErrorNotInitialized;  // prepare() sizes the grid
  if (grid > total) grid = total;
  hipLaunchKernelGGL((k2d_f64p<MF, R, SH, KF, BY, BX, NW, ST>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in,
                     out, bfrag, g);
  return hipGetLastError();
}

#define CS_V2DP(ID, MF, R, SH, KF, BY, BX, NW, ST)                                             \
  Variant {                                                                                    \
    ID, MF, R, SH, KF, BY, BX, NW, Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::LDS,                   \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::PX, ST, &launch_k2dp<MF, R, SH, KF, BY, BX, NW, ST>, \
        &prepare_k2dp<MF, R, SH, KF, BY, BX, NW, ST>,                                          \
        "k2d_f64p<mf" #MF ",r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ",st" #ST ">",   \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::CS                                                \
  }

#define CS_V2DU(ID, MF, R, SH, KF, BY, BX, NW, ST)                                               \
  Variant {                                                                                        \
    ID, MF, R, SH, KF, BY, BX, NW, Cfg2D<MF, R, SH, KF, BY, BX, NW, ST, kAligned>::LDS,            \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST, kAligned>::PX, ST, &launch_k2du<MF, R, SH, KF, BY, BX, NW, ST>, \
        &prepare_k2du<MF, R, SH, KF, BY, BX, NW, ST>,                                               \
        "k2d_f64<mf" #MF ",r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ",st" #ST ",aligned128>", \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST, kAligned>::CS                                         \
  }

#define CS_V2D(ID, MF, R, SH, KF, BY, BX, NW, ST)                                            \
  Variant {                                                                                  \
    ID, MF, R, SH, KF, BY, BX, NW, Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::LDS,                 \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::PX, ST, &launch_k2d<MF, R, SH, KF, BY, BX, NW, ST>, \
        &prepare_k2d<MF, R, SH, KF, BY, BX, NW, ST>,                                         \
        "k2d_f64<mf" #MF ",r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ",st" #ST ">",   \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::CS                                              \
  }

// Variant ids: ST * 10000 + R * 1000 + (MF == 4) * 100 + (SH == star) * 10 + KF % 10, plus
// 100000 for alternative tile shapes and depths 6..9, 200000 for depths 10..19, and
// 500000 more for the persistent form.  The host
// takes the first entry that matches (MFMA shape, radius, shape class, fusion depth), so
// preferred variants come first.
static const Variant kVariants[] = {
    // single-step launches (a halo that is not zero, and the reference's pre-fused
    // weights): the output staged in LDS and written row by row (ST 2) wins at every radius
    CS_V2D(711101, 4, 1, SHAPE_BOX, 1, 64, 64, 8, 2),
    CS_V2D(711111, 4, 1, SHAPE_STAR, 1, 64, 64, 8, 2),
    CS_V2D(713101, 4, 3, SHAPE_BOX, 1, 64, 64, 8, 2),
    CS_V2D(713111, 4, 3, SHAPE_STAR, 1, 64, 64, 8, 2),
    // register-staged, v_mfma_f64_4x4x4 on 8 x 8 blocks (default family)
    CS_V2D(11101, 4, 1, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(11102, 4, 1, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(11103, 4, 1, SHAPE_BOX, 3, 64, 64, 8, 1),
    CS_V2D(11104, 4, 1, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_V2D(111106, 4, 1, SHAPE_BOX, 6, 64, 64, 8, 1),
    CS_V2D(111108, 4, 1, SHAPE_BOX, 8, 64, 64, 8, 1),
    CS_V2D(211100, 4, 1, SHAPE_BOX, 10, 64, 64, 8, 1),
    CS_V2D(211102, 4, 1, SHAPE_BOX, 12, 64, 64, 8, 1),
    CS_V2D(11111, 4, 1, SHAPE_STAR, 1, 64, 64, 8, 1),
    CS_V2D(11112, 4, 1, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_V2D(11113, 4, 1, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_V2D(11114, 4, 1, SHAPE_STAR, 4, 64, 64, 8, 1),
    CS_V2D(111116, 4, 1, SHAPE_STAR, 6, 64, 64, 8, 1),
    CS_V2D(111118, 4, 1, SHAPE_STAR, 8, 64, 64, 8, 1),
    CS_V2D(12101, 4, 2, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(12102, 4, 2, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(12103, 4, 2, SHAPE_BOX, 3, 64, 64, 8, 1),
    CS_V2D(12104, 4, 2, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_V2D(12111, 4, 2, SHAPE_STAR, 1, 64, 64, 8, 1),
    CS_V2D(12112, 4, 2, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_V2D(12113, 4, 2, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_V2D(12114, 4, 2, SHAPE_STAR, 4, 64, 64, 8, 1),
    CS_V2D(13101, 4, 3, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(13102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(13103, 4, 3, SHAPE_BOX, 3, 64, 64, 8, 1),
    CS_V2D(13104, 4, 3, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_V2D(13111, 4, 3, SHAPE_STAR, 1, 64, 64, 8, 1),
    CS_V2D(13112, 4, 3, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_V2D(13113, 4, 3, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_V2D(13114, 4, 3, SHAPE_STAR, 4, 64, 64, 8, 1),
    // register-staged, v_mfma_f64_16x16x4 on 16 x 16 blocks
    CS_V2D(11001, 16, 1, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(11002, 16, 1, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(13001, 16, 3, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(13011, 16, 3, SHAPE_STAR, 1, 64, 64, 8, 1),
    // two LDS buffers (ping-pong), for comparison
    CS_V2D(1102, 4, 1, SHAPE_BOX, 2, 64, 64, 8, 0),
    CS_V2D(1104, 4, 1, SHAPE_BOX, 4, 64, 64, 8, 0),
    CS_V2D(3102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 0),
    // other tile shapes and wave counts
    CS_V2D(111104, 4, 1, SHAPE_BOX, 4, 128, 64, 8, 1),
    CS_V2D(111204, 4, 1, SHAPE_BOX, 4, 64, 64, 4, 1),
    CS_V2D(111208, 4, 1, SHAPE_BOX, 8, 64, 64, 4, 1),
    CS_V2D(113202, 4, 3, SHAPE_BOX, 2, 64, 64, 4, 1),
    // last step staged through LDS, coalesced output rows (ids 700000 + R * 1000 + ...)
    CS_V2D(711102, 4, 1, SHAPE_BOX, 2, 64, 64, 8, 2),
    CS_V2D(711104, 4, 1, SHAPE_BOX, 4, 64, 64, 8, 2),
    CS_V2D(711106, 4, 1, SHAPE_BOX, 6, 64, 64, 8, 2),
    CS_V2D(711108, 4, 1, SHAPE_BOX, 8, 64, 64, 8, 2),
    CS_V2D(711114, 4, 1, SHAPE_STAR, 4, 64, 64, 8, 2),
    CS_V2D(711118, 4, 1, SHAPE_STAR, 8, 64, 64, 8, 2),
    CS_V2D(713102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 2),
    CS_V2D(713112, 4, 3, SHAPE_STAR, 2, 64, 64, 8, 2),
    CS_V2D(713113, 4, 3, SHAPE_STAR, 3, 64, 64, 8, 2),
    CS_V2D(713114, 4, 3, SHAPE_STAR, 4, 64, 64, 8, 2),
    // persistent form (ids 500000 + the id of the same configuration)
    CS_V2DP(511101, 4, 1, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2DP(511102, 4, 1, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2DP(511104, 4, 1, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_V2DP(611106, 4, 1, SHAPE_BOX, 6, 64, 64, 8, 1),
    CS_V2DP(611108, 4, 1, SHAPE_BOX, 8, 64, 64, 8, 1),
    CS_V2DP(611204, 4, 1, SHAPE_BOX, 4, 64, 64, 4, 1),
    CS_V2DP(511114, 4, 1, SHAPE_STAR, 4, 64, 64, 8, 1),
    CS_V2DP(611116, 4, 1, SHAPE_STAR, 6, 64, 64, 8, 1),
    CS_V2DP(611118, 4, 1, SHAPE_STAR, 8, 64, 64, 8, 1),
    CS_V2DP(512102, 4, 2, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2DP(513101, 4, 3, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2DP(513102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2DP(513103, 4, 3, SHAPE_BOX, 3, 64, 64, 8, 1),
    CS_V2DP(513112, 4, 3, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_V2DP(513113, 4, 3, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_V2DP(513114, 4, 3, SHAPE_STAR, 4, 64, 64, 8, 1),
    // the --breakdown ladder's rung 3: single-step kernels with 128-byte aligned LDS rows
    // (ids 930000 + R * 1000 + 100 + shape * 10 + KF; listed last so that they are never
    // the first match)
    CS_V2DU(933101, 4, 3, SHAPE_BOX, 1, 64, 64, 8, 1), CS_V2DU(931101, 4, 1, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2DU(933111, 4, 3, SHAPE_STAR, 1, 64, 64, 8, 1), CS_V2DU(931111, 4, 1, SHAPE_STAR, 1, 64, 64, 8, 1),
    // the production depths with 128-byte aligned rows (id + 9000000), for the autotuner
    CS_V2DU(9711106, 4, 1, SHAPE_BOX, 6, 64, 64, 8, 2), CS_V2DU(9711108, 4, 1, SHAPE_BOX, 8, 64, 64, 8, 2),
    CS_V2DU(9111106, 4, 1, SHAPE_BOX, 6, 64, 64, 8, 1), CS_V2DU(9013103, 4, 3, SHAPE_BOX, 3, 64, 64, 8, 1),
    CS_V2DU(9013102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 1), CS_V2DU(9013113, 4, 3, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_V2DU(9713113, 4, 3, SHAPE_STAR, 3, 64, 64, 8, 2), CS_V2DU(9713101, 4, 3, SHAPE_BOX, 1, 64, 64, 8, 2),
    CS_V2DU(9713111, 4, 3, SHAPE_STAR, 1, 64, 64, 8, 2), CS_V2DU(9711101, 4, 1, SHAPE_BOX, 1, 64, 64, 8, 2),
    CS_V2DU(9711111, 4, 1, SHAPE_STAR, 1, 64, 64, 8, 2),
};

const Variant *variants_2d_f64(int *count) {
  *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
}

void build_bfrag_2d_f64(int mf, int R, int shape, const double *w, double *bfrag) {
  const int D = 2 * R + 1;
  if (mf == 1 || mf == 2) {  // the ladder's FMA kernels read the dense weights
    const int n = bfrag_ksteps_2d_f64(mf, R, shape) * 64;
    for (int i = 0; i < n; ++i) bfrag[i] = i < D * D ? w[i] : 0.0;
    return;
  }
  if (mf == 16) {
    const KTable kt = make_ktable(R, shape);
    for (int ks = 0; ks < kt.n; ++ks)
      for (int lane = 0; lane < 64; ++lane) {
        const int kk = lane >> 4, n = lane & 15;
        const int dx = kt.c0[ks] + kk - n;
        double v = 0.0;
        if (dx >= -R && dx <= R) v = w[(kt.dy[ks] + R) * D + (dx + R)];
        bfrag[ks * 64 + lane] = v;
      }
  } else {
    // dilated table: k-step ks, lane group k = lane >> 4 holds tap entry t; output
    // column n = lane & 3 is the cell moved by v_n = (n >> 1, n & 1), so B[t][n] = W[t - v_n]
    const DTable dt = make_dtable(R, shape);
    auto put = [&](int ks, int k, int ty, int tx, int live) {
      for (int lane = 16 * k; lane < 16 * k + 16; ++lane) {
        const int n = lane & 3;
        const int dy = ty - (n >> 1), dx = tx - (n & 1);
        double v = 0.0;
        if (live && dy >= -R && dy <= R && dx >= -R && dx <= R) v = w[(dy + R) * D + (dx + R)];
        bfrag[ks * 64 + lane] = v;
      }
    };
    for (int ps = 0; ps < dt.np; ++ps)
      for (int k = 0; k < 4; ++k)
        for (int h = 0; h < 2; ++h) put(2 * ps + h, k, dt.pty[ps][k], dt.ptx[ps][k] + h, dt.plive[ps][k]);
    for (int ss = 0; ss < dt.ns; ++ss)
      for (int k = 0; k < 4; ++k) put(2 * dt.np + ss, k, dt.sty[ss][k], dt.stx[ss][k], dt.slive[ss][k]);
  }
}

int bfrag_ksteps_2d_f64(int mf, int R, int shape) {
  if (mf == 1 || mf == 2) return ((2 * R + 1) * (2 * R + 1) + 63) / 64;
  return mf == 16 ? make_ktable(R, shape).n : dtable_ksteps(make_dtable(R, shape));
}

}  // namespace cs
