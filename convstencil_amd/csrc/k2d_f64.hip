// k2d_f64.hip -- plain FP64 2-D ConvStencil kernels for gfx950 (FP64 MFMA).
//
// Replaces kernel2d (reference src/2d/gpu.cu:30-70) and its launchers
// gpu_box_2d1r / gpu_box_2d3r (src/2d/gpu.cu:245-425).  What changes for CDNA4:
//   * the A operand is the input tile itself, read straight from LDS (see
//     cs_ktable.h), so the reference's lookup-table scatter into two stencil2row
//     copies is gone and each input element is stored in LDS once;
//   * B is a banded Toeplitz slice of the weights, held in registers for the
//     whole kernel, one f64 per lane per k-step;
//   * two MFMA shapes: v_mfma_f64_16x16x4 (16x16 output blocks) and
//     v_mfma_f64_4x4x4 (4 blocks of 4x4 = 4x16 output strips), which run at the
//     same flop rate on gfx950 but waste fewer flops on the band's zeros;
//   * the time loop is fused inside the kernel: a workgroup loads a tile with a
//     halo of KF*R cells, advances it KF steps in LDS (ping-pong between two LDS
//     buffers), keeping the domain's halo cells fixed after every step, and writes
//     only the final BY x BX block.  This equals KF separate steps exactly in exact
//     arithmetic -- unlike the reference's pre-composed K^3 weights, which differ
//     from stepping in the outer (KF-1)*R ring of the domain;
//   * workgroups are mapped to tiles so that each XCD (own L2) sweeps a contiguous
//     band of tile rows, keeping vertical halo re-reads inside one L2.
#include "cs_internal.h"

namespace cs {

template <int MF, int R, int SH, int KF, int BY, int BX, int NW>
struct Cfg2D {
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  // 16x16x4 tables read up to OVR columns past a block's last column; 4x4x4 tables never do
  static constexpr int OVR = MF == 16 ? ktable_overrun(make_ktable(R, SH)) : 0;
  static constexpr int PX0 = EX + OVR + 1;
  static constexpr int PX = (PX0 % 2 == 0) ? PX0 + 1 : PX0;  // odd pitch spreads rows over banks
  static constexpr int NBUF = KF > 1 ? 2 : 1;
  static constexpr int BUF = EY * PX;
  static constexpr size_t LDS = (size_t)NBUF * BUF * sizeof(double);
  static constexpr int BH = MF == 16 ? 16 : 4;  // output block height
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
};

__device__ __forceinline__ int64_t xcd_tile(int64_t wg, int64_t ntiles) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
  // contiguous tile range [x * per, (x + 1) * per).
  const int64_t per = (ntiles + 7) / 8;
  return (wg % 8) * per + wg / 8;
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW>
__global__ __launch_bounds__(NW * 64) void k2d_f64(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW>;
  constexpr KTable kt = MF == 16 ? make_ktable(R, SH) : make_ftable(R, SH);
  constexpr int NKS = MF == 16 ? kt.n : kt.n / 4;
  constexpr int BH = C::BH;
  extern __shared__ double lds[];

  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;  // uniform over the workgroup

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t pitch = g.ext[2];
  const int64_t ny = g.n[1], nx = g.n[2];
  // interior coordinates of tile cell (0, 0)
  const int64_t Y0 = (tile / g.tiles[2]) * BY - KF * R;
  const int64_t X0 = (tile % g.tiles[2]) * BX - KF * R;
  const int64_t ay0 = g.org[1] + Y0, ax0 = g.org[2] + X0;

  double b[NKS];
#pragma unroll
  for (int ks = 0; ks < NKS; ++ks) b[ks] = bfrag[ks * 64 + lane];

  // Per-lane LDS offsets of the A operand relative to the block's (y, x) cell.
  int off[NKS];
  if constexpr (MF == 16) {
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) off[ks] = ((lane & 15) + kt.dy[ks]) * C::PX + kt.c0[ks] + (lane >> 4);
  } else {
    const int k = lane >> 4, blk4 = (lane >> 2) & 3, i = lane & 3;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int f = 4 * ks + k;
      off[ks] = (i + kt.dy[f]) * C::PX + 4 * blk4 + kt.c0[f];
    }
  }

  // ---- load the tile (and its halo) into both LDS buffers ----
  for (int idx = tid; idx < C::BUF; idx += NW * 64) {
    const int r = idx / C::PX, c = idx - r * C::PX;
    const int64_t ay = ay0 + r, ax = ax0 + c;
    double v = 0.0;
    if (c < C::EX && ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < pitch) v = in[ay * pitch + ax];
    lds[idx] = v;
    if constexpr (KF > 1) lds[C::BUF + idx] = v;
  }
  __syncthreads();

  const bool full = Y0 >= 0 && X0 >= 0 && Y0 + C::EY <= ny && X0 + C::EX <= nx;

#pragma unroll
  for (int s = 0; s < KF; ++s) {
    const double *src = lds + (s & 1) * C::BUF;
    double *dst = lds + ((s + 1) & 1) * C::BUF;
    const int H = BY + 2 * R * (KF - 1 - s), W = BX + 2 * R * (KF - 1 - s);
    const int base = R * (s + 1);
    const int nby = (H + BH - 1) / BH, nbx = (W + 15) / 16;
    for (int blk = wave; blk < nby * nbx; blk += NW) {
      const int ba = blk / nbx, bb = blk - ba * nbx;
      const int y = base + min(BH * ba, H - BH);
      const int x = base + min(16 * bb, W - 16);
      const double *p = src + y * C::PX + x;
      if constexpr (MF == 16) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(p[off[ks]], b[ks], acc, 0, 0, 0);
        const int tx = x + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ty = y + 4 * v + (lane >> 4);
          const bool inside = full || (Y0 + ty >= 0 && Y0 + ty < ny && X0 + tx >= 0 && X0 + tx < nx);
          if (inside) {
            if (s < KF - 1)
              dst[ty * C::PX + tx] = acc[v];
            else
              out[(ay0 + ty) * pitch + ax0 + tx] = acc[v];
          }
        }
      } else {
        double acc = 0.0;
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) acc = __builtin_amdgcn_mfma_f64_4x4x4f64(p[off[ks]], b[ks], acc, 0, 0, 0);
        const int ty = y + (lane >> 4), tx = x + (lane & 15);
        const bool inside = full || (Y0 + ty >= 0 && Y0 + ty < ny && X0 + tx >= 0 && X0 + tx < nx);
        if (inside) {
          if (s < KF - 1)
            dst[ty * C::PX + tx] = acc;
          else
            out[(ay0 + ty) * pitch + ax0 + tx] = acc;
        }
      }
    }
    if (s < KF - 1) __syncthreads();
  }
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW>
static hipError_t prepare_k2d() {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW>;
  return hipFuncSetAttribute((const void *)k2d_f64<MF, R, SH, KF, BY, BX, NW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW>
static hipError_t launch_k2d(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW>;
  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k2d_f64<MF, R, SH, KF, BY, BX, NW>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in, out,
                     bfrag, g);
  return hipGetLastError();
}

#define CS_V2D(ID, MF, R, SH, KF, BY, BX, NW)                                                                  \
  Variant {                                                                                                  \
    ID, MF, R, SH, KF, BY, BX, NW, Cfg2D<MF, R, SH, KF, BY, BX, NW>::LDS, Cfg2D<MF, R, SH, KF, BY, BX, NW>::PX, \
        &launch_k2d<MF, R, SH, KF, BY, BX, NW>,                                                           \
        &prepare_k2d<MF, R, SH, KF, BY, BX, NW>,                                                        \
        "k2d_f64<mf" #MF ",r" #R ",sh" #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ">"                              \
  }

static const Variant kVariants[] = {
    // radius 1 (box2d1r 9-point, star2d1r 5-point)
    CS_V2D(1001, 16, 1, SHAPE_BOX, 1, 64, 64, 8),
    CS_V2D(1002, 16, 1, SHAPE_BOX, 2, 64, 64, 8),
    CS_V2D(1101, 4, 1, SHAPE_BOX, 1, 64, 64, 8),
    CS_V2D(1102, 4, 1, SHAPE_BOX, 2, 64, 64, 8),
    CS_V2D(1103, 4, 1, SHAPE_BOX, 3, 64, 64, 8),
    CS_V2D(1104, 4, 1, SHAPE_BOX, 4, 64, 64, 8),
    CS_V2D(1111, 4, 1, SHAPE_STAR, 1, 64, 64, 8),
    CS_V2D(1112, 4, 1, SHAPE_STAR, 2, 64, 64, 8),
    // radius 2
    CS_V2D(2101, 4, 2, SHAPE_BOX, 1, 64, 64, 8),
    CS_V2D(2111, 4, 2, SHAPE_STAR, 1, 64, 64, 8),
    // radius 3 (box2d3r 49-point, star2d3r 13-point, the reference's pre-fused 7x7)
    CS_V2D(3001, 16, 3, SHAPE_BOX, 1, 64, 64, 8),
    CS_V2D(3101, 4, 3, SHAPE_BOX, 1, 64, 64, 8),
    CS_V2D(3102, 4, 3, SHAPE_BOX, 2, 64, 64, 8),
    CS_V2D(3011, 16, 3, SHAPE_STAR, 1, 64, 64, 8),
    CS_V2D(3111, 4, 3, SHAPE_STAR, 1, 64, 64, 8),
    CS_V2D(3112, 4, 3, SHAPE_STAR, 2, 64, 64, 8),
};

const Variant *variants_2d_f64(int *count) {
  *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
}

void build_bfrag_2d_f64(int mf, int R, int shape, const double *w, double *bfrag) {
  const int D = 2 * R + 1;
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
    const KTable ft = make_ftable(R, shape);
    for (int ks = 0; ks < ft.n / 4; ++ks)
      for (int lane = 0; lane < 64; ++lane) {
        const int f = 4 * ks + (lane >> 4), j = lane & 3;
        const int dx = ft.c0[f] - j;
        double v = 0.0;
        if (ft.live[f] && dx >= -R && dx <= R) v = w[(ft.dy[f] + R) * D + (dx + R)];
        bfrag[ks * 64 + lane] = v;
      }
  }
}

int bfrag_ksteps_2d_f64(int mf, int R, int shape) {
  return mf == 16 ? make_ktable(R, shape).n : make_ftable(R, shape).n / 4;
}

}  // namespace cs
