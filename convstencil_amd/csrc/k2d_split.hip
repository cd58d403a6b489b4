// k2d_split.hip -- split-precision 2-D ConvStencil kernels for gfx950.
//
// The state stays FP64 in HBM.  Inside a fused launch each tile value u and each
// weight w is split into slices of a narrow type,
//   BF16x3 : u = u0 + u1 + u2 (bf16),                  v_mfma_f32_16x16x32_bf16
//   FP16x2 : u = u0 + u1 * 2^-11 (fp16, rescaled),     v_mfma_f32_16x16x32_f16
//   FP32x2 : u = u0 + u1 (fp32 hi / lo),               v_mfma_f32_16x16x4f32
// and the stencil is a sum of slice products accumulated in FP32, one
// accumulator per precision level s + t (levels above the kept ones are dropped:
// BF16x3 keeps s + t <= 2, i.e. 6 of 9 products; the 2-slice forms keep 3 of 4).
// Products of two 8-bit (bf16) or 11-bit (fp16) significands are exact in FP32,
// so the only FP32 rounding is in the accumulation.  After every fused step the
// levels are recombined in FP64 (the rescaling is a power of two), halo cells are
// left untouched, and the FP64 value is re-split for the next step; the last step
// writes FP64 to HBM.
//
// GEMM form (output block of 16 rows x 16 columns): A rows are output columns and
// B columns are output rows, with the Toeplitz along y:
//   A[i][(dx, rr)] = u(y - R + rr, x + i + dx),   B[(dx, rr)][n] = w(rr - R - n, dx).
// Slice planes are stored column-major in LDS, so the eight consecutive rows a
// lane feeds to one MFMA are one aligned 16-byte read.  Row storage is offset so
// that every block's first input row y - R starts an 8-row chunk; 3 chunks (24
// rows) then cover the 16 + 2R rows a block needs (R <= 4).
#include "cs_internal.h"

namespace cs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

struct TrBF16 {
  using T = __bf16;
  using V = bf16x8;
  static constexpr int CK = 8, KI = 32, NSU = 3, NSW = 3, NL = 3;
  static constexpr double SCALE = 1.0;  // slice s carries weight SCALE^-s
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
struct TrF16 {
  using T = _Float16;
  using V = f16x8;
  static constexpr int CK = 8, KI = 32, NSU = 2, NSW = 2, NL = 2;
  static constexpr double SCALE = 2048.0;
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
struct TrF32 {
  using T = float;
  using V = float;
  static constexpr int CK = 1, KI = 4, NSU = 2, NSW = 2, NL = 2;
  static constexpr double SCALE = 1.0;
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// Splits v into NS slices of Tr::T; slice s is scaled by SCALE^s before rounding.
template <class Tr>
__host__ __device__ inline void split_value(double v, typename Tr::T *s) {
  double rem = v, sc = 1.0;
#pragma unroll
  for (int k = 0; k < Tr::NSU; ++k) {
    const typename Tr::T t = (typename Tr::T)(float)(rem * sc);
    s[k] = t;
    rem -= (double)(float)t / sc;
    sc *= Tr::SCALE;
  }
}

template <class Tr, int R, int KF, int BY, int BX, int NW>
struct CfgS {
  using T = typename Tr::T;
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  // rows entries per dx, and the storage offset that aligns every block's first input row
  static constexpr int RPD = Tr::CK == 8 ? 24 : ((16 + 2 * R + 3) / 4) * 4;
  static constexpr int ALIGN = Tr::CK == 8 ? 8 : 1;
  static constexpr int TOP0 = ((R - KF * R) % 8 + 8) % 8;  // TOP == R - KF*R (mod 8)
  static constexpr int TOP = Tr::CK == 8 ? TOP0 + 8 : 8;
  static constexpr int PR0 = TOP + EY + RPD - 2 * R - 1;   // rows incl. the last block's overrun
  static constexpr int PR = (PR0 + 7) / 8 * 8;
  static constexpr int PLANE = EX * PR;                     // elements per slice plane
  static constexpr int NSU = Tr::NSU;
  static constexpr int BUF = NSU * PLANE;                   // one buffer: all slices
  static constexpr size_t LDS = 2 * (size_t)BUF * sizeof(T);
  static constexpr int NENT = (2 * R + 1) * RPD;            // K entries (dx, rr)
  static constexpr int NINS = (NENT + Tr::KI - 1) / Tr::KI; // MFMA instructions per product
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(R <= 4, "3 chunks of 8 rows cover 16 + 2R rows only for R <= 4");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
  // K entry e -> (dx, rr); padding entries map to (0, 0) with weight 0
  static constexpr int ent_dx(int e) { return e < NENT ? e / RPD - R : 0; }
  static constexpr int ent_rr(int e) { return e < NENT ? e % RPD : 0; }
};

template <class Tr, int R, int KF, int BY, int BX, int NW>
__global__ __launch_bounds__(NW * 64) void k2d_split(const double *__restrict__ in, double *__restrict__ out,
                                                     const typename Tr::T *__restrict__ bfrag, Geo g) {
  using C = CfgS<Tr, R, KF, BY, BX, NW>;
  using T = typename Tr::T;
  using V = typename Tr::V;
  constexpr int NINS = C::NINS, NSU = Tr::NSU, NSW = Tr::NSW, NL = Tr::NL, CK = Tr::CK;
  extern __shared__ __align__(16) unsigned char smem[];
  T *lds = reinterpret_cast<T *>(smem);

  const int64_t ntiles = g.tiles[1] * g.tiles[2];
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t pitch = g.ext[2], ny = g.n[1], nx = g.n[2];
  const int64_t Y0 = (tile / g.tiles[2]) * BY - KF * R;
  const int64_t X0 = (tile % g.tiles[2]) * BX - KF * R;
  const int64_t ay0 = g.org[1] + Y0, ax0 = g.org[2] + X0;
  const bool full = Y0 >= 0 && X0 >= 0 && Y0 + C::EY <= ny && X0 + C::EX <= nx;

  // B fragments: b[t][m] for weight slice t and instruction m
  V b[NSW][NINS];
#pragma unroll
  for (int t = 0; t < NSW; ++t)
#pragma unroll
    for (int m = 0; m < NINS; ++m) {
      const T *src = bfrag + ((size_t)(t * NINS + m) * 64 + lane) * CK;
      if constexpr (CK == 8)
        b[t][m] = *reinterpret_cast<const V *>(src);
      else
        b[t][m] = *src;
    }
  // per-lane A offsets (elements, within a slice plane) relative to the block's
  // (x, y - R + TOP) origin: lane group q = lane >> 4 takes K entries of chunk q
  int off[NINS];
  const int ai = lane & 15, q = lane >> 4;
#pragma unroll
  for (int m = 0; m < NINS; ++m) {
    const int e = m * Tr::KI + q * CK;
    off[m] = (ai + C::ent_dx(e)) * C::PR + C::ent_rr(e);
  }

  // ---- load: FP64 tile -> slices in both buffers (halo cells stay as loaded) ----
  const bool all_in = ay0 >= 0 && ay0 + C::EY <= g.ext[1] && ax0 >= 0 && ax0 + C::EX <= pitch;
  for (int idx = tid; idx < C::PLANE; idx += NW * 64) {
    const int c = idx / C::PR, sr = idx - c * C::PR;  // column-major: consecutive threads walk rows
    const int r = sr - C::TOP;
    double v = 0.0;
    if (r >= 0 && r < C::EY) {
      const int64_t ay = ay0 + r, ax = ax0 + c;
      if (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < pitch)) v = in[ay * pitch + ax];
    }
    T s[NSU];
    split_value<Tr>(v, s);
#pragma unroll
    for (int k = 0; k < NSU; ++k) {
      lds[k * C::PLANE + idx] = s[k];
      lds[C::BUF + k * C::PLANE + idx] = s[k];
    }
  }
  __syncthreads();

  for (int st = 0; st < KF; ++st) {
    const T *src = lds + (st & 1) * C::BUF;
    T *dst = lds + ((st + 1) & 1) * C::BUF;
    const bool last = st == KF - 1;
    const int lo = R * (st + 1), hi = C::EY - R * (st + 1);  // valid rows (and columns) of this step
    const int W = C::EX - 2 * R * (st + 1);
    // first block row: the largest y <= lo with y == KF*R (mod 8) (alignment of y - R + TOP)
    const int y0 = lo - (((lo - KF * R) % 8) + 8) % 8;
    const int nby = (hi - y0 + 15) / 16, nbx = (W + 15) / 16;
    for (int blk = wave; blk < nby * nbx; blk += NW) {
      const int ba = blk / nbx, bb = blk - ba * nbx;
      const int y = y0 + 16 * ba;
      const int x = lo + min(16 * bb, W - 16);
      const int ubase = x * C::PR + (y - R + C::TOP);
      f4 acc[NL];
#pragma unroll
      for (int L = 0; L < NL; ++L) acc[L] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int m = 0; m < NINS; ++m) {
#pragma unroll
        for (int s = 0; s < NSU; ++s) {
          V a;
          if constexpr (CK == 8)
            a = *reinterpret_cast<const V *>(src + s * C::PLANE + ubase + off[m]);
          else
            a = src[s * C::PLANE + ubase + off[m]];
#pragma unroll
          for (int t = 0; t < NSW; ++t)
            if (s + t < NL) acc[s + t] = Tr::mma(a, b[t][m], acc[s + t]);
        }
      }
      // recombine in FP64: lane holds output column x + 4 (lane >> 4) + v, row y + (lane & 15)
      const int tyy = y + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int txx = x + 4 * (lane >> 4) + v;
        double val = 0.0, sc = 1.0;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          val += (double)acc[L][v] / sc;
          sc *= Tr::SCALE;
        }
        const bool in_region = tyy >= lo && tyy < hi;
        const bool inside = full || (Y0 + tyy >= 0 && Y0 + tyy < ny && X0 + txx >= 0 && X0 + txx < nx);
        if (!(in_region && inside)) continue;
        if (last) {
          out[(ay0 + tyy) * pitch + ax0 + txx] = val;
        } else {
          T sl[NSU];
          split_value<Tr>(val, sl);
#pragma unroll
          for (int k = 0; k < NSU; ++k) dst[k * C::PLANE + txx * C::PR + tyy + C::TOP] = sl[k];
        }
      }
    }
    if (!last) __syncthreads();
  }
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

// Variant::mf carries the precision (1 = BF16x3, 2 = FP32x2, 3 = FP16x2); st = 0.
#define CS_VS(ID, PREC, TR, R, KF, BY, BX, NW)                                                                     \
  Variant {                                                                                                        \
    ID, PREC, R, SHAPE_BOX, KF, BY, BX, NW, CfgS<TR, R, KF, BY, BX, NW>::LDS, CfgS<TR, R, KF, BY, BX, NW>::PR, 0,  \
        &launch_ks<TR, R, KF, BY, BX, NW>, &prepare_ks<TR, R, KF, BY, BX, NW>,                                     \
        "k2d_split<" #TR ",r" #R ",kf" #KF "," #BY "x" #BX ",w" #NW ">"                                            \
  }

static const Variant kVariantsSplit[] = {
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

size_t build_bfrag_split(int prec, int R, const double *w, void *out, size_t cap) {
  unsigned char *o = reinterpret_cast<unsigned char *>(out);
  if (prec == 1) return build_split<TrBF16>(R, w, o, cap);
  if (prec == 2) return build_split<TrF32>(R, w, o, cap);
  if (prec == 3) return build_split<TrF16>(R, w, o, cap);
  return 0;
}

}  // namespace cs
