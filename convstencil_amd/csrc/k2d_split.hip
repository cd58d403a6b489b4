// k2d_split.hip -- split-precision 2-D ConvStencil kernels: the dilated form
// (k2d_split2_body.h, BF16x3 and FP16x2, preferred) and the Toeplitz form
// (k2d_split_body.h, also FP32x2).
#include "cs_wcomp.h"
#include "k2d_split2_body.h"

namespace cs {

template <class Tr, int R, int KF, int BY, int BX, int NW>
__global__ __launch_bounds__(NW * 64) void k2d_split(const double *__restrict__ in, double *__restrict__ out,
                                                     const typename Tr::T *__restrict__ bfrag, Geo g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;
  tile_split<Tr, R, KF, BY, BX, NW>(in, out, bfrag, g, tile, smem, threadIdx.x);
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW, int OS>
__global__ __launch_bounds__(NW * 64) void k2d_split2(const double *__restrict__ in, double *__restrict__ out,
                                                      const typename Tr::T *__restrict__ bfrag, Geo g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;
  tile_split2<Tr, R, SH, KF, BY, BX, NW, 64, OS>(in, out, bfrag, g, tile, smem, threadIdx.x);
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW, int OS>
static hipError_t prepare_ks2() {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW, OS>;
  return hipFuncSetAttribute((const void *)k2d_split2<Tr, R, SH, KF, BY, BX, NW, OS>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW, int OS>
static hipError_t launch_ks2(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW, OS>;
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k2d_split2<Tr, R, SH, KF, BY, BX, NW, OS>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s,
                     in, out, reinterpret_cast<const typename Tr::T *>(bfrag), g);
  return hipGetLastError();
}

// Persistent dilated form: a fixed grid (a multiple of 8) walks the tiles in the
// XCD-aware order; while a tile is computed, one load per thread touches every cache line
// of the next tile (cs_internal.h, prefetch_tile), so its load comes from L2.
__device__ __forceinline__ int64_t next_tile_pos_s(int64_t k, int64_t G, int64_t total, int64_t ntiles) {
  while (k < total && (k % 8) * ((ntiles + 7) / 8) + k / 8 >= ntiles) k += G;
  return k;
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW>
__global__ __launch_bounds__(NW * 64) void k2d_split2p(const double *__restrict__ in, double *__restrict__ out,
                                                       const typename Tr::T *__restrict__ bfrag, Geo g) {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW>;
  using T = typename Tr::T;
  using V = typename Tr::V;
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t per = (ntiles + 7) / 8, total = per * 8, G = gridDim.x;
  int64_t k = next_tile_pos_s(blockIdx.x, G, total, ntiles);
  if (k >= total) return;
  SplitCtx t;
  V w[C::NSW][C::NINS];
  int off[C::NINS];
  split_lane_setup<C>(bfrag, threadIdx.x, t, w, off);
  split_tile_setup<C>(g, (k % 8) * per + k / 8, t, out);
  for (;;) {
    // opaque copies of the thread and wave indices: keep the compiler from hoisting the
    // per-element and per-block addresses out of the tile loop
    int tid = t.tid, wv = t.wave;
    asm volatile("" : "+v"(tid), "+s"(wv));
    t.tid = tid;
    t.wave = wv;
    if (t.full)
      load_tile_split_full<C>(in, t, lds);
    else
      load_tile_split<C, 64>(in, g, t, lds);
    lds_barrier();
    const int64_t kn = next_tile_pos_s(k + G, G, total, ntiles);
    const bool more = kn < total;
    SplitCtx tn = t;
    double pf[Prefetch<C>::N];
    if (more) {
      split_tile_setup<C>(g, (kn % 8) * per + kn / 8, tn, out);
      prefetch_tile<C>(in, g, tn, pf);  // in flight while this tile is computed
    }
    run_split_steps<C>(t, w, off, lds);
    if (!more) break;
    prefetch_consume<C>(pf);
    lds_barrier();  // every wave is done with this tile's LDS before the next tile is stored
    k = kn;
    t = tn;
  }
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW>
static int &ks2p_grid() {
  static int grid = 0;
  return grid;
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW>
static hipError_t prepare_ks2p() {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW>;
  const void *fn = (const void *)k2d_split2p<Tr, R, SH, KF, BY, BX, NW>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
  if (e != hipSuccess) return e;
  int per_cu = 0, dev = 0, cus = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NW * 64, C::LDS);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  ks2p_grid<Tr, R, SH, KF, BY, BX, NW>() = ((per_cu > 0 ? per_cu : 1) * cus) / 8 * 8;
  return hipSuccess;
}

template <class Tr, int R, int SH, int KF, int BY, int BX, int NW>
static hipError_t launch_ks2p(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW>;
  const int64_t total = ((g.tiles[1] * g.tiles[2] + 7) / 8) * 8;
  int64_t grid = ks2p_grid<Tr, R, SH, KF, BY, BX, NW>();
  if (grid <= 0) return hipErrorNotInitialized;  // prepare() sizes the grid
  if (grid > total) grid = total;
  hipLaunchKernelGGL((k2d_split2p<Tr, R, SH, KF, BY, BX, NW>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in,
                     out, reinterpret_cast<const typename Tr::T *>(bfrag), g);
  return hipGetLastError();
}

template <class Tr, int R, int KF, int BY, int BX, int NW>
static hipError_t prepare_ks() {
  using C = CfgS<Tr, R, KF, BY, BX, NW>;
  return hipFuncSetAttribute((const void *)k2d_split<Tr, R, KF, BY, BX, NW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <class Tr, int R, int KF, int BY, int BX, int NW>
static hipError_t launch_ks(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = CfgS<Tr, R, KF, BY, BX, NW>;
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k2d_split<Tr, R, KF, BY, BX, NW>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in, out,
                     reinterpret_cast<const typename Tr::T *>(bfrag), g);
  return hipGetLastError();
}

// Variant::mf carries the precision (1 = BF16x3, 2 = FP32x2, 3 = FP16x2, 4 = FP16 with
// three weight slices); st = 0.
#define CS_VS(ID, PREC, TR, R, KF, BY, BX, NW)                                                                     \
  Variant {                                                                                                        \
    ID, PREC, R, SHAPE_BOX, KF, BY, BX, NW, CfgS<TR, R, KF, BY, BX, NW>::LDS, CfgS<TR, R, KF, BY, BX, NW>::PR, 0,  \
        &launch_ks<TR, R, KF, BY, BX, NW>, &prepare_ks<TR, R, KF, BY, BX, NW>,                                     \
        "k2d_split<" #TR ",r" #R ",kf" #KF "," #BY "x" #BX ",w" #NW ">"                                            \
  }

// Dilated form: Variant::st = 2.  Ids 60000 + R * 1000 + precision * 100 + (star) * 10 + KF;
// 700000 + that: staged output (OS = 1); 800000 + that: 128 x 64 tiles; 900000 + that:
// 128 x 128 tiles with 16 waves.
#define CS_VS2O(ID, PREC, TR, R, SH, KF, BY, BX, NW, OS)                                                \
  Variant {                                                                                             \
    ID, PREC, R, SH, KF, BY, BX, NW, CfgS2<TR, R, SH, KF, BY, BX, NW, OS>::LDS,                         \
        CfgS2<TR, R, SH, KF, BY, BX, NW, OS>::PX, 2, &launch_ks2<TR, R, SH, KF, BY, BX, NW, OS>,       \
        &prepare_ks2<TR, R, SH, KF, BY, BX, NW, OS>,                                                    \
        "k2d_split2<" #TR ",r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ",os" #OS ">",             \
        CfgS2<TR, R, SH, KF, BY, BX, NW, OS>::CS                                                        \
  }
#define CS_VS2(ID, PREC, TR, R, SH, KF, BY, BX, NW) CS_VS2O(ID, PREC, TR, R, SH, KF, BY, BX, NW, 0)

#define CS_VS2P(ID, PREC, TR, R, SH, KF, BY, BX, NW)                                                    \
  Variant {                                                                                             \
    ID, PREC, R, SH, KF, BY, BX, NW, CfgS2<TR, R, SH, KF, BY, BX, NW>::LDS,                             \
        CfgS2<TR, R, SH, KF, BY, BX, NW>::PX, 2, &launch_ks2p<TR, R, SH, KF, BY, BX, NW>,              \
        &prepare_ks2p<TR, R, SH, KF, BY, BX, NW>,                                                       \
        "k2d_split2p<" #TR ",r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ">",                      \
        CfgS2<TR, R, SH, KF, BY, BX, NW>::CS                                                            \
  }

static const Variant kVariantsSplit[] = {
    CS_VS2(61101, 1, TrBF16, 1, SHAPE_BOX, 1, 64, 64, 8),
    CS_VS2(61102, 1, TrBF16, 1, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2(61103, 1, TrBF16, 1, SHAPE_BOX, 3, 64, 64, 8),
    CS_VS2(61104, 1, TrBF16, 1, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2(61111, 1, TrBF16, 1, SHAPE_STAR, 1, 64, 64, 8),
    CS_VS2(61112, 1, TrBF16, 1, SHAPE_STAR, 2, 64, 64, 8),
    CS_VS2(61114, 1, TrBF16, 1, SHAPE_STAR, 4, 64, 64, 8),
    CS_VS2(63101, 1, TrBF16, 3, SHAPE_BOX, 1, 64, 64, 8),
    CS_VS2(63102, 1, TrBF16, 3, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2(63103, 1, TrBF16, 3, SHAPE_BOX, 3, 64, 64, 8),
    CS_VS2(63104, 1, TrBF16, 3, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2(63105, 1, TrBF16, 3, SHAPE_BOX, 5, 64, 64, 8),
    CS_VS2(63106, 1, TrBF16, 3, SHAPE_BOX, 6, 64, 64, 8),
    CS_VS2(63111, 1, TrBF16, 3, SHAPE_STAR, 1, 64, 64, 8),
    CS_VS2(63112, 1, TrBF16, 3, SHAPE_STAR, 2, 64, 64, 8),
    CS_VS2(63113, 1, TrBF16, 3, SHAPE_STAR, 3, 64, 64, 8),
    CS_VS2(63114, 1, TrBF16, 3, SHAPE_STAR, 4, 64, 64, 8),
    CS_VS2(61301, 3, TrF16, 1, SHAPE_BOX, 1, 64, 64, 8),
    CS_VS2(61302, 3, TrF16, 1, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2(61304, 3, TrF16, 1, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2(63301, 3, TrF16, 3, SHAPE_BOX, 1, 64, 64, 8),
    CS_VS2(63302, 3, TrF16, 3, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2(63303, 3, TrF16, 3, SHAPE_BOX, 3, 64, 64, 8),
    CS_VS2(63304, 3, TrF16, 3, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2(63305, 3, TrF16, 3, SHAPE_BOX, 5, 64, 64, 8),
    CS_VS2(63306, 3, TrF16, 3, SHAPE_BOX, 6, 64, 64, 8),
    CS_VS2(63311, 3, TrF16, 3, SHAPE_STAR, 1, 64, 64, 8),
    CS_VS2(63307, 3, TrF16, 3, SHAPE_BOX, 7, 64, 64, 8),
    CS_VS2(63308, 3, TrF16, 3, SHAPE_BOX, 8, 64, 64, 8),
    CS_VS2(63313, 3, TrF16, 3, SHAPE_STAR, 3, 64, 64, 8),
    CS_VS2(63314, 3, TrF16, 3, SHAPE_STAR, 4, 64, 64, 8),
    CS_VS2(61401, 4, TrF16W3, 1, SHAPE_BOX, 1, 64, 64, 8),
    CS_VS2(61402, 4, TrF16W3, 1, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2(61404, 4, TrF16W3, 1, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2(63401, 4, TrF16W3, 3, SHAPE_BOX, 1, 64, 64, 8),
    CS_VS2(63402, 4, TrF16W3, 3, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2(63403, 4, TrF16W3, 3, SHAPE_BOX, 3, 64, 64, 8),
    CS_VS2(63404, 4, TrF16W3, 3, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2(63405, 4, TrF16W3, 3, SHAPE_BOX, 5, 64, 64, 8),
    CS_VS2(63406, 4, TrF16W3, 3, SHAPE_BOX, 6, 64, 64, 8),
    CS_VS2(63411, 4, TrF16W3, 3, SHAPE_STAR, 1, 64, 64, 8),
    CS_VS2(63412, 4, TrF16W3, 3, SHAPE_STAR, 2, 64, 64, 8),
    CS_VS2(63413, 4, TrF16W3, 3, SHAPE_STAR, 3, 64, 64, 8),
    CS_VS2(63414, 4, TrF16W3, 3, SHAPE_STAR, 4, 64, 64, 8),
    CS_VS2(63312, 3, TrF16, 3, SHAPE_STAR, 2, 64, 64, 8),
    // staged FP64 output of the last step (ids 700000 + the id of the same configuration)
    CS_VS2O(761102, 1, TrBF16, 1, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_VS2O(761104, 1, TrBF16, 1, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_VS2O(763102, 1, TrBF16, 3, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_VS2O(763104, 1, TrBF16, 3, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_VS2O(763106, 1, TrBF16, 3, SHAPE_BOX, 6, 64, 64, 8, 1),
    // 128 x 64 tiles (one workgroup per CU, less halo work per output)
    CS_VS2O(863104, 1, TrBF16, 3, SHAPE_BOX, 4, 128, 64, 8, 0),
    CS_VS2O(863406, 4, TrF16W3, 3, SHAPE_BOX, 6, 128, 64, 8, 0),
    // 128 x 128 tiles, 16 waves (one workgroup per CU; two-slice state leaves room for
    // deep fusion)
    CS_VS2O(963104, 1, TrBF16, 3, SHAPE_BOX, 4, 128, 128, 16, 0),
    CS_VS2O(963404, 4, TrF16W3, 3, SHAPE_BOX, 4, 128, 128, 16, 0),
    CS_VS2O(963406, 4, TrF16W3, 3, SHAPE_BOX, 6, 128, 128, 16, 0),
    CS_VS2O(963408, 4, TrF16W3, 3, SHAPE_BOX, 8, 128, 128, 16, 0),
    CS_VS2O(963414, 4, TrF16W3, 3, SHAPE_STAR, 4, 128, 128, 16, 0),
    CS_VS2O(863306, 3, TrF16, 3, SHAPE_BOX, 6, 128, 64, 8, 0),
    CS_VS2O(963306, 3, TrF16, 3, SHAPE_BOX, 6, 128, 128, 16, 0),
    CS_VS2O(963308, 3, TrF16, 3, SHAPE_BOX, 8, 128, 128, 16, 0),
    // FP16x2 kf6 launch shapes (ids 1000000 + shape index * 100000 + the base id)
    CS_VS2O(1063306, 3, TrF16, 3, SHAPE_BOX, 6, 64, 64, 4, 0),
    CS_VS2O(1163306, 3, TrF16, 3, SHAPE_BOX, 6, 64, 64, 16, 0),
    CS_VS2O(1263306, 3, TrF16, 3, SHAPE_BOX, 6, 48, 64, 8, 0),
    CS_VS2O(1363306, 3, TrF16, 3, SHAPE_BOX, 6, 64, 48, 8, 0),
    CS_VS2O(1463306, 3, TrF16, 3, SHAPE_BOX, 6, 80, 64, 8, 0),
    CS_VS2O(1563306, 3, TrF16, 3, SHAPE_BOX, 6, 64, 80, 8, 0),
    CS_VS2O(1463305, 3, TrF16, 3, SHAPE_BOX, 5, 80, 64, 8, 0),
    CS_VS2O(1663306, 3, TrF16, 3, SHAPE_BOX, 6, 80, 64, 10, 0),
    CS_VS2O(1763306, 3, TrF16, 3, SHAPE_BOX, 6, 80, 64, 12, 0),
    CS_VS2O(1463406, 4, TrF16W3, 3, SHAPE_BOX, 6, 80, 64, 8, 0),
    CS_VS2O(1463314, 3, TrF16, 3, SHAPE_STAR, 4, 80, 64, 8, 0),
    CS_VS2O(1463316, 3, TrF16, 3, SHAPE_STAR, 6, 80, 64, 8, 0),
    CS_VS2O(763112, 1, TrBF16, 3, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_VS2O(763113, 1, TrBF16, 3, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_VS2O(763304, 3, TrF16, 3, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_VS2O(763404, 4, TrF16W3, 3, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_VS2O(763406, 4, TrF16W3, 3, SHAPE_BOX, 6, 64, 64, 8, 1),
    // persistent dilated form (ids 500000 + the id of the same configuration)
    CS_VS2P(561102, 1, TrBF16, 1, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2P(561104, 1, TrBF16, 1, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2P(563102, 1, TrBF16, 3, SHAPE_BOX, 2, 64, 64, 8),
    CS_VS2P(563103, 1, TrBF16, 3, SHAPE_BOX, 3, 64, 64, 8),
    CS_VS2P(563104, 1, TrBF16, 3, SHAPE_BOX, 4, 64, 64, 8),
    CS_VS2P(563112, 1, TrBF16, 3, SHAPE_STAR, 2, 64, 64, 8),
    CS_VS2P(563113, 1, TrBF16, 3, SHAPE_STAR, 3, 64, 64, 8),
    CS_VS2P(563303, 3, TrF16, 3, SHAPE_BOX, 3, 64, 64, 8),
    CS_VS2P(563304, 3, TrF16, 3, SHAPE_BOX, 4, 64, 64, 8),
    // Toeplitz form

    CS_VS(51101, 1, TrBF16, 1, 1, 64, 64, 8), CS_VS(51102, 1, TrBF16, 1, 2, 64, 64, 8),
    CS_VS(51104, 1, TrBF16, 1, 4, 64, 64, 8), CS_VS(53101, 1, TrBF16, 3, 1, 64, 64, 8),
    CS_VS(53102, 1, TrBF16, 3, 2, 64, 64, 8), CS_VS(53301, 3, TrF16, 3, 1, 64, 64, 8),
    CS_VS(53302, 3, TrF16, 3, 2, 64, 64, 8),  CS_VS(53303, 3, TrF16, 3, 3, 64, 64, 8),
    CS_VS(51301, 3, TrF16, 1, 1, 64, 64, 8),  CS_VS(51302, 3, TrF16, 1, 2, 64, 64, 8),
    CS_VS(51304, 3, TrF16, 1, 4, 64, 64, 8),  CS_VS(53201, 2, TrF32, 3, 1, 64, 64, 8),
    CS_VS(53202, 2, TrF32, 3, 2, 64, 64, 8),  CS_VS(51201, 2, TrF32, 1, 1, 64, 64, 8),
    CS_VS(51202, 2, TrF32, 1, 2, 64, 64, 8),
};

const Variant *variants_2d_split(int *count) {
  *count = (int)(sizeof(kVariantsSplit) / sizeof(kVariantsSplit[0]));
  return kVariantsSplit;
}

// B fragments for precision prec (1 BF16x3, 2 FP32x2, 3 FP16x2), written as raw slice
// elements into `out` (bytes); returns the byte count, or 0 if cap is too small.
template <class Tr>
static size_t build_split(int R, const double *w, unsigned char *out, size_t cap) {
  using T = typename Tr::T;
  const int rpd = Tr::CK == 8 ? 24 : ((16 + 2 * R + 3) / 4) * 4;
  const int nent = (2 * R + 1) * rpd;
  const int nins = (nent + Tr::KI - 1) / Tr::KI;
  const size_t bytes = (size_t)Tr::NSW * nins * 64 * Tr::CK * sizeof(T);
  if (bytes > cap) return 0;
  T *o = reinterpret_cast<T *>(out);
  const int D = 2 * R + 1;
  for (int t = 0; t < Tr::NSW; ++t)
    for (int m = 0; m < nins; ++m)
      for (int lane = 0; lane < 64; ++lane)
        for (int e = 0; e < Tr::CK; ++e) {
          const int q = lane >> 4, n = lane & 15;
          const int ent = m * Tr::KI + q * Tr::CK + e;
          double v = 0.0;
          if (ent < nent) {
            const int dx = ent / rpd - R, rr = ent % rpd;
            const int dy = rr - R - n;
            if (dy >= -R && dy <= R) v = w[(dy + R) * D + (dx + R)];
          }
          T sl[Tr::NSU];
          split_value<Tr>(v, sl);
          o[((size_t)(t * nins + m) * 64 + lane) * Tr::CK + e] = sl[t];
        }
  return bytes;
}

// A fragments of the dilated form: weight slice q, instruction i, lane l, element j is
// slice q of W[t - v_m] with m = l & 15, chunk 4 i + (l >> 4) and t = (ty, tx0 + j).
template <class Tr>
static size_t build_split2(int R, int shape, const double *w, unsigned char *out, size_t cap) {
  using T = typename Tr::T;
  const CTable ct = make_ctable(R, shape);
  const int nins = ct.n / 4;
  const size_t bytes = (size_t)Tr::NSW * nins * 64 * 8 * sizeof(T);
  if (bytes > cap || R > 3) return 0;
  T *o = reinterpret_cast<T *>(out);
  const int D = 2 * R + 1;
  T tap[3][49] = {};  // slices of every tap
  for (int t = 0; t < D * D; ++t) {
    T sl[3];
    split_weight<Tr>(w[t], sl);
    for (int q = 0; q < Tr::NSW; ++q) tap[q][t] = sl[q];
  }
  if constexpr (Tr::NSW == Tr::NSU) {
    // BF16x3 and FP16x2: the last weight slice is moved by ulps until the weights' errors
    // cancel in the mass and the first moments (cs_wcomp.h); FP16 w3's third slice leaves
    // errors near 2^-36 and needs none
    double fixed[49];
    int off[49][3];
    const double last_scale = Tr::NSW == 3 ? Tr::SCALE * Tr::SCALE : Tr::SCALE;
    for (int t = 0; t < D * D; ++t) {
      fixed[t] = (double)(float)tap[0][t] + (Tr::NSW == 3 ? (double)(float)tap[1][t] / Tr::SCALE : 0.0);
      off[t][0] = 0;
      off[t][1] = t / D - R;
      off[t][2] = t % D - R;
    }
    compensate_low_slice<T>(D * D, w, off, fixed, tap[Tr::NSW - 1], last_scale);
  }
  for (int q = 0; q < Tr::NSW; ++q)
    for (int i = 0; i < nins; ++i)
      for (int lane = 0; lane < 64; ++lane)
        for (int j = 0; j < 8; ++j) {
          const int m = lane & 15, c = 4 * i + (lane >> 4);
          const int dy = ct.ty[c] - (m >> 3), dx = ct.tx[c] + j - (m & 7);
          T v = (T)0.0f;
          if (ct.live[c] && dy >= -R && dy <= R && dx >= -R && dx <= R) v = tap[q][(dy + R) * D + (dx + R)];
          o[((size_t)(q * nins + i) * 64 + lane) * 8 + j] = v;
        }
  return bytes;
}

size_t build_bfrag_split2(int prec, int R, int shape, const double *w, void *out, size_t cap) {
  unsigned char *o = reinterpret_cast<unsigned char *>(out);
  if (prec == 1) return build_split2<TrBF16>(R, shape, w, o, cap);
  if (prec == 3) return build_split2<TrF16>(R, shape, w, o, cap);
  if (prec == 4) return build_split2<TrF16W3>(R, shape, w, o, cap);
  return 0;
}

size_t build_bfrag_split(int prec, int R, const double *w, void *out, size_t cap) {
  unsigned char *o = reinterpret_cast<unsigned char *>(out);
  if (prec == 1) return build_split<TrBF16>(R, w, o, cap);
  if (prec == 2) return build_split<TrF32>(R, w, o, cap);
  if (prec == 3) return build_split<TrF16>(R, w, o, cap);
  return 0;
}

}  // namespace cs
