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
//     halo of KF*R cells, advances it KF steps in LDS, keeping the domain's halo
//     cells fixed after every step, and writes only the final BY x BX block.  This
//     equals KF separate steps exactly in exact arithmetic -- unlike the
//     reference's pre-composed K^3 weights, which differ from stepping in the outer
//     (KF-1)*R ring of the domain;
//   * two staging modes for the fused steps: ST = 0 ping-pongs between two LDS
//     buffers; ST = 1 keeps one buffer and stages each step's results in registers
//     (every wave holds all of its blocks' accumulators, i.e. many independent
//     MFMA chains), writing them back between two barriers.  ST = 1 halves LDS, so
//     several workgroups share a CU and one's HBM loads overlap the others' MFMAs;
//   * workgroups are mapped to tiles so that each XCD (own L2) sweeps a contiguous
//     band of tile rows, keeping vertical halo re-reads inside one L2.
#include "cs_internal.h"

namespace cs {

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
struct Cfg2D {
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  // 16x16x4 tables read up to OVR columns past a block's last column; 4x4x4 tables never do
  static constexpr int OVR = MF == 16 ? ktable_overrun(make_ktable(R, SH)) : 0;
  static constexpr int PX0 = EX + OVR + 1;
  static constexpr int PX = (PX0 % 2 == 0) ? PX0 + 1 : PX0;  // odd pitch spreads rows over banks
  static constexpr int NBUF = (KF > 1 && ST == 0) ? 2 : 1;
  static constexpr int BUF = EY * PX;
  static constexpr size_t LDS = (size_t)NBUF * BUF * sizeof(double);
  static constexpr int BH = MF == 16 ? 16 : 4;  // output block height
  static constexpr KTable KT = MF == 16 ? make_ktable(R, SH) : make_ftable(R, SH);
  static constexpr int NKS = MF == 16 ? KT.n : KT.n / 4;
  // region of fused step s, its block grid, and the most blocks one wave handles in a step
  static constexpr int H(int s) { return BY + 2 * R * (KF - 1 - s); }
  static constexpr int W(int s) { return BX + 2 * R * (KF - 1 - s); }
  static constexpr int NBY(int s) { return (H(s) + BH - 1) / BH; }
  static constexpr int NBX(int s) { return (W(s) + 15) / 16; }
  static constexpr int PER = (NBY(0) * NBX(0) + NW - 1) / NW;
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
};

__device__ __forceinline__ int64_t xcd_tile(int64_t wg, int64_t ntiles) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
  // contiguous tile range [x * per, (x + 1) * per).
  const int64_t per = (ntiles + 7) / 8;
  return (wg % 8) * per + wg / 8;
}

// MFMA result of one block: 4 f64 per lane (16x16x4) or 1 (4x4x4).
template <int MF>
struct Acc;
template <>
struct Acc<16> {
  d4 v;
};
template <>
struct Acc<4> {
  double v;
};

template <int MF, int NKS>
__device__ __forceinline__ Acc<MF> block_chain(const double *p, const int *off, const double *b) {
  Acc<MF> a;
  if constexpr (MF == 16) {
    a.v = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) a.v = __builtin_amdgcn_mfma_f64_16x16x4f64(p[off[ks]], b[ks], a.v, 0, 0, 0);
  } else {
    a.v = 0.0;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) a.v = __builtin_amdgcn_mfma_f64_4x4x4f64(p[off[ks]], b[ks], a.v, 0, 0, 0);
  }
  return a;
}

// Writes a block's results to LDS (tile coordinates) or to the output grid.
// Cells outside the domain interior are skipped: they keep their halo value.
template <int MF, int PX, bool GLOBAL>
__device__ __forceinline__ void put_block(const Acc<MF> &a, int y, int x, int lane, bool full, int64_t Y0, int64_t X0,
                                          int64_t ny, int64_t nx, double *lds_dst, double *out, int64_t ay0,
                                          int64_t ax0, int64_t pitch) {
  if constexpr (MF == 16) {
    const int tx = x + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ty = y + 4 * v + (lane >> 4);
      if (full || (Y0 + ty >= 0 && Y0 + ty < ny && X0 + tx >= 0 && X0 + tx < nx)) {
        if constexpr (GLOBAL)
          out[(ay0 + ty) * pitch + ax0 + tx] = a.v[v];
        else
          lds_dst[ty * PX + tx] = a.v[v];
      }
    }
  } else {
    const int ty = y + (lane >> 4), tx = x + (lane & 15);
    if (full || (Y0 + ty >= 0 && Y0 + ty < ny && X0 + tx >= 0 && X0 + tx < nx)) {
      if constexpr (GLOBAL)
        out[(ay0 + ty) * pitch + ax0 + tx] = a.v;
      else
        lds_dst[ty * PX + tx] = a.v;
    }
  }
}

template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST>
__global__ __launch_bounds__(NW * 64) void k2d_f64(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>;
  constexpr KTable kt = C::KT;
  constexpr int NKS = C::NKS;
  constexpr int BH = C::BH;
  constexpr int PX = C::PX;
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
  const bool full = Y0 >= 0 && X0 >= 0 && Y0 + C::EY <= ny && X0 + C::EX <= nx;

  double b[NKS];
#pragma unroll
  for (int ks = 0; ks < NKS; ++ks) b[ks] = bfrag[ks * 64 + lane];

  // Per-lane LDS offsets of the A operand relative to the block's (y, x) cell.
  int off[NKS];
  if constexpr (MF == 16) {
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) off[ks] = ((lane & 15) + kt.dy[ks]) * PX + kt.c0[ks] + (lane >> 4);
  } else {
    const int k = lane >> 4, blk4 = (lane >> 2) & 3, i = lane & 3;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int f = 4 * ks + k;
      off[ks] = (i + kt.dy[f]) * PX + 4 * blk4 + kt.c0[f];
    }
  }

  // ---- load the tile (and its halo); cells outside the array and pad columns are 0 ----
  const bool all_in = ay0 >= 0 && ay0 + C::EY <= g.ext[1] && ax0 >= 0 && ax0 + C::EX <= pitch;
  for (int idx = tid; idx < C::BUF; idx += NW * 64) {
    const int r = idx / PX, c = idx - r * PX;
    const int64_t ay = ay0 + r, ax = ax0 + c;
    double v = 0.0;
    if (c < C::EX && (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < pitch))) v = in[ay * pitch + ax];
    lds[idx] = v;
    if constexpr (C::NBUF == 2) lds[C::BUF + idx] = v;
  }
  __syncthreads();

#pragma unroll
  for (int s = 0; s < KF; ++s) {
    const int H = C::H(s), W = C::W(s);
    const int base = R * (s + 1);
    const int nbx = C::NBX(s), nb = C::NBY(s) * C::NBX(s);
    const bool last = s == KF - 1;
    if constexpr (ST == 0) {
      const double *src = lds + (s & 1) * C::BUF;
      double *dst = lds + ((s + 1) & 1) * C::BUF;
      for (int blk = wave; blk < nb; blk += NW) {
        const int ba = blk / nbx, bb = blk - ba * nbx;
        const int y = base + min(BH * ba, H - BH), x = base + min(16 * bb, W - 16);
        const Acc<MF> a = block_chain<MF, NKS>(src + y * PX + x, off, b);
        if (last)
          put_block<MF, PX, true>(a, y, x, lane, full, Y0, X0, ny, nx, nullptr, out, ay0, ax0, pitch);
        else
          put_block<MF, PX, false>(a, y, x, lane, full, Y0, X0, ny, nx, dst, nullptr, 0, 0, 0);
      }
      if (!last) __syncthreads();
    } else {
      Acc<MF> acc[C::PER];
#pragma unroll
      for (int j = 0; j < C::PER; ++j) {
        const int blk = j * NW + wave;
        if (blk < nb) {
          const int ba = blk / nbx, bb = blk - ba * nbx;
          const int y = base + min(BH * ba, H - BH), x = base + min(16 * bb, W - 16);
          acc[j] = block_chain<MF, NKS>(lds + y * PX + x, off, b);
        }
      }
      if (!last) __syncthreads();  // every wave has read this step's input
#pragma unroll
      for (int j = 0; j < C::PER; ++j) {
        const int blk = j * NW + wave;
        if (blk < nb) {
          const int ba = blk / nbx, bb = blk - ba * nbx;
          const int y = base + min(BH * ba, H - BH), x = base + min(16 * bb, W - 16);
          if (last)
            put_block<MF, PX, true>(acc[j], y, x, lane, full, Y0, X0, ny, nx, nullptr, out, ay0, ax0, pitch);
          else
            put_block<MF, PX, false>(acc[j], y, x, lane, full, Y0, X0, ny, nx, lds, nullptr, 0, 0, 0);
        }
      }
      if (!last) __syncthreads();  // this step's output is complete before it is read
    }
  }
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

#define CS_V2D(ID, MF, R, SH, KF, BY, BX, NW, ST)                                            \
  Variant {                                                                                  \
    ID, MF, R, SH, KF, BY, BX, NW, Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::LDS,                 \
        Cfg2D<MF, R, SH, KF, BY, BX, NW, ST>::PX, ST, &launch_k2d<MF, R, SH, KF, BY, BX, NW, ST>, \
        &prepare_k2d<MF, R, SH, KF, BY, BX, NW, ST>,                                         \
        "k2d_f64<mf" #MF ",r" #R "," #SH ",kf" #KF "," #BY "x" #BX ",w" #NW ",st" #ST ">"    \
  }

// Variant ids: ST * 10000 + R * 1000 + (MF == 4) * 100 + (SH == star) * 10 + KF, plus
// 100000 for alternative tile shapes.  The host takes the first entry that matches
// (MFMA shape, radius, shape class, fusion depth), so preferred variants come first.
static const Variant kVariants[] = {
    // register-staged, v_mfma_f64_4x4x4 (default family)
    CS_V2D(11101, 4, 1, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(11102, 4, 1, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(11103, 4, 1, SHAPE_BOX, 3, 64, 64, 8, 1),
    CS_V2D(11104, 4, 1, SHAPE_BOX, 4, 64, 64, 8, 1),
    CS_V2D(11111, 4, 1, SHAPE_STAR, 1, 64, 64, 8, 1),
    CS_V2D(11112, 4, 1, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_V2D(11113, 4, 1, SHAPE_STAR, 3, 64, 64, 8, 1),
    CS_V2D(11114, 4, 1, SHAPE_STAR, 4, 64, 64, 8, 1),
    CS_V2D(12101, 4, 2, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(12102, 4, 2, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(12111, 4, 2, SHAPE_STAR, 1, 64, 64, 8, 1),
    CS_V2D(12112, 4, 2, SHAPE_STAR, 2, 64, 64, 8, 1),
    CS_V2D(13101, 4, 3, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(13102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(13111, 4, 3, SHAPE_STAR, 1, 64, 64, 8, 1),
    CS_V2D(13112, 4, 3, SHAPE_STAR, 2, 64, 64, 8, 1),
    // register-staged, v_mfma_f64_16x16x4
    CS_V2D(11001, 16, 1, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(11002, 16, 1, SHAPE_BOX, 2, 64, 64, 8, 1),
    CS_V2D(13001, 16, 3, SHAPE_BOX, 1, 64, 64, 8, 1),
    CS_V2D(13011, 16, 3, SHAPE_STAR, 1, 64, 64, 8, 1),
    // two LDS buffers (ping-pong), for comparison
    CS_V2D(1102, 4, 1, SHAPE_BOX, 2, 64, 64, 8, 0),
    CS_V2D(1104, 4, 1, SHAPE_BOX, 4, 64, 64, 8, 0),
    CS_V2D(3102, 4, 3, SHAPE_BOX, 2, 64, 64, 8, 0),
    // other tile shapes and depths
    CS_V2D(111104, 4, 1, SHAPE_BOX, 4, 128, 64, 8, 1),
    CS_V2D(111106, 4, 1, SHAPE_BOX, 6, 64, 64, 8, 1),
    CS_V2D(111108, 4, 1, SHAPE_BOX, 8, 64, 64, 8, 1),
    CS_V2D(111204, 4, 1, SHAPE_BOX, 4, 64, 64, 4, 1),
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
